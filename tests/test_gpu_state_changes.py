"""The residual after every way the committed state can change.

Bar: after each transition the residual equals the oracle's for the state the
oracle holds (within 1e-12 of max(|f_int|, |f_lin|), the color accumulation
order), the same call repeated gives the same bits, and the committed plastic
strain equals the oracle's bit for bit:
  * zero state;
  * a committing residual (update_mem), then update_mem off;
  * ipf_compute_sigma with update_mem (the separate-step path), then off;
  * ipf_set_state with a state of the oracle, then back to zeros.
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_residual_follows_every_state_change(gpu):
    import torch
    m = wb.make_mesh((12, 9, 2, 1.2))
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    s = sm.sizes()
    neq, npts = s["neq"], s["npts"]
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    d = m.material()
    mat = (d["lambda"], d["mu"], d["phi"], d["psi"], d["cohesion"])
    ctx = sm.context()
    flin = ix["rhs_linear"]
    xy = m.nodes()
    u1 = wb.lame_displacement(xy, wb.normal_stress(10), scale=1.5)
    u2 = wb.lame_displacement(xy, wb.normal_stress(10), scale=1.8)

    def gpu_residual(u, commit=False):
        r = torch.full((neq,), float("nan"), dtype=torch.float64, device=gpu)
        ctx.set_update_mem(commit)
        ctx.residual(torch.as_tensor(u, device=gpu), r)
        ctx.set_update_mem(False)
        ctx.synchronize()
        return r.cpu().numpy()

    def check(u, ep, al):
        r_o, *_ = O.residual(B, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], neq, ix["weight"], *mat, u,
                             f_lin=flin, plastic_strain6=ep, alpha=al)
        tol = 1e-12 * max(np.abs(r_o - flin).max(), np.abs(flin).max())
        r1 = gpu_residual(u)
        r2 = gpu_residual(u)
        assert np.array_equal(r1, r2)
        assert np.abs(r1 - r_o).max() <= tol

    zero_ep, zero_al = np.zeros(6 * npts), np.zeros(npts)
    check(u2, zero_ep, zero_al)
    # commit through the fused residual
    gpu_residual(u1, commit=True)
    _, ep1, al1, mt1, _ = O.residual(B, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], neq, ix["weight"], *mat,
                                     u1, f_lin=flin, update_mem=True)
    assert (mt1 != 0).any() and (mt1 == 0).any(), "the load should make some, not all, points plastic"
    ep_g, al_g, _, _ = ctx.get_state()
    assert np.array_equal(ep_g, ep1)
    check(u2, ep1, al1)
    # commit through the separate steps (ipf_compute_sigma with update_mem)
    eps = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ws = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ctx.multiply(torch.as_tensor(u2, device=gpu), eps)
    ctx.set_update_mem(True)
    ctx.compute_sigma(eps, ws)
    ctx.set_update_mem(False)
    ctx.synchronize()
    _, ep2, al2, _, _ = O.residual(B, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], neq, ix["weight"], *mat,
                                   u2, f_lin=flin, plastic_strain6=ep1, alpha=al1, update_mem=True)
    assert np.array_equal(ctx.get_state()[0], ep2)
    check(u1, ep2, al2)
    # ipf_set_state: a state from the host, then back to zeros
    ctx.set_state(ep1, al1)
    check(u2, ep1, al1)
    ctx.set_state(zero_ep, zero_al)
    check(u2, zero_ep, zero_al)
