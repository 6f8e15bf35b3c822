"""GPU parity of the HIP path against the CPU oracle (oracle/), through the C ABI.

Bars (stated per test):
  * strains (B.u) and per-point branch flags / mtype: bit-exact;
  * per-point w*sigma and committed plastic strain: bit-exact (same operation
    order, no fused multiply-add on either side);
  * residual and B^T products: the GPU accumulates colors in a different
    order than the oracle's slab reduction, so they agree to rounding:
    |r_gpu - r_oracle|_inf <= 1e-12 * max(|f_int|_inf, |f_lin|_inf) -- far
    inside the 1e-10 relative bar.
"""
import math

import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL_SUM = 1e-12


def _setup(mesh_tuple, sn=None):
    m = wb.make_mesh(mesh_tuple)
    if sn is not None:
        m.set_normal_stress(sn, 0.0)
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    return m, sm, B, ix


def _mat(m):
    d = m.material()
    return d["lambda"], d["mu"], d["phi"], d["psi"], d["cohesion"]


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def _np(t):
    return t.detach().cpu().numpy()


MESHES = [(6, 5, 1, 1.2), (40, 40, 2, 1.1), (5, 4, 3, 1.3)]


@pytest.mark.parametrize("mesh", MESHES, ids=lambda t: f"{t[0]}x{t[1]}-q{t[2]}")
@pytest.mark.parametrize("scale", [1.0, 3.0, 12.0])
def test_steps_match_oracle(gpu, mesh, scale):
    import torch
    m, sm, B, ix = _setup(mesh)
    lam, mu, phi, psi, c = _mat(m)
    ctx = sm.context()
    u = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=scale)
    npts = B.rows // 3
    # Multiply: bit-exact strains
    eps_o = O.multiply(B, ix["dof_indexes"], u)
    du = _t(u, gpu)
    eps_g = torch.zeros(B.rows, dtype=torch.float64, device=gpu)
    ctx.multiply(du, eps_g)
    ctx.synchronize()
    assert np.array_equal(_np(eps_g), eps_o)
    # ComputeSigma with commit: bit-exact stresses, flags and plastic strain
    ws_o, ep_o, al_o, mt_o, br_o = O.compute_sigma(eps_o, ix["weight"], lam, mu, phi, psi, c, update_mem=True)
    ws_g = torch.zeros(B.rows, dtype=torch.float64, device=gpu)
    ctx.set_update_mem(True)
    ctx.compute_sigma(eps_g, ws_g)
    ctx.set_update_mem(False)
    ctx.synchronize()
    assert np.array_equal(_np(ws_g), ws_o)
    ep_g, al_g, mt_g, br_g = ctx.get_state()
    assert np.array_equal(mt_g, mt_o)
    assert np.array_equal(br_g, br_o)
    assert np.array_equal(ep_g, ep_o)
    assert np.array_equal(al_g, al_o)
    if scale >= 3.0:
        assert (mt_o != 0).any(), "expected plastic points at this load"
    # MultiplyTranspose: rounding-level agreement
    res_o = O.multiply_transpose(B, ix["color_indexes"], sm.sizes()["ncolor"], m.info()["neq"], ws_o)
    res_g = torch.zeros(m.info()["neq"], dtype=torch.float64, device=gpu)
    ctx.multiply_transpose(ws_g, res_g)
    ctx.synchronize()
    scale_f = np.abs(res_o).max()
    assert np.abs(_np(res_g) - res_o).max() <= RTOL_SUM * scale_f


@pytest.mark.parametrize("mesh", MESHES, ids=lambda t: f"{t[0]}x{t[1]}-q{t[2]}")
def test_residual_and_commit_match_oracle(gpu, mesh):
    """Fused residual (Assemble(rhs)) at two consecutive load states, with the
    state committed in between (AcceptPseudoTimeStepSolution)."""
    import torch
    m, sm, B, ix = _setup(mesh)
    lam, mu, phi, psi, c = _mat(m)
    ctx = sm.context()
    neq = m.info()["neq"]
    ncolor = sm.sizes()["ncolor"]
    flin = ix["rhs_linear"]
    xy = m.nodes()
    ep = None
    al = None
    for k, scale in enumerate((6.0, 9.0)):
        u = wb.lame_displacement(xy, wb.normal_stress(10), scale=scale)
        r_o, ep_n, al_n, mt_o, br_o = O.residual(B, ix["dof_indexes"], ix["color_indexes"], ncolor, neq,
                                                 ix["weight"], lam, mu, phi, psi, c, u, f_lin=flin,
                                                 plastic_strain6=ep, alpha=al, update_mem=True)
        du = _t(u, gpu)
        dr = torch.zeros(neq, dtype=torch.float64, device=gpu)
        ctx.set_update_mem(True)
        ctx.residual(du, dr)
        ctx.set_update_mem(False)
        ctx.synchronize()
        fint = r_o - flin
        tol = RTOL_SUM * max(np.abs(fint).max(), np.abs(flin).max())
        assert np.abs(_np(dr) - r_o).max() <= tol
        ep_g, al_g, mt_g, br_g = ctx.get_state()
        assert np.array_equal(mt_g, mt_o)
        assert np.array_equal(br_g, br_o)
        assert np.array_equal(ep_g, ep_n)
        assert np.array_equal(al_g, al_n)
        ep, al = ep_n, al_n
        # host-pointer entry point of the struct matrix gives the same (uncommitted) residual
        r_h = sm.assemble_rhs(u)
        r_o2, *_ = O.residual(B, ix["dof_indexes"], ix["color_indexes"], ncolor, neq, ix["weight"], lam, mu, phi,
                              psi, c, u, f_lin=flin, plastic_strain6=ep, alpha=al, update_mem=False)
        assert np.abs(r_h - r_o2).max() <= tol


def test_multiply_vector_matches_oracle(gpu):
    import torch
    m, sm, B, ix = _setup((7, 6, 2, 1.1))
    ctx = sm.context()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(B.cols)
    y_o = O.multiply_vector(B, x, 0)
    y_g = torch.zeros(B.rows, dtype=torch.float64, device=gpu)
    ctx.multiply_vector(_t(x, gpu), y_g, 0)
    ctx.synchronize()
    assert np.array_equal(_np(y_g), y_o)
    s = rng.standard_normal(B.rows)
    z_o = O.multiply_vector(B, s, 1)
    z_g = torch.zeros(B.cols, dtype=torch.float64, device=gpu)
    ctx.multiply_vector(_t(s, gpu), z_g, 1)
    ctx.synchronize()
    assert np.array_equal(_np(z_g), z_o)


def test_compute_sigma_dep_matches_oracle(gpu):
    import torch
    m, sm, B, ix = _setup((10, 8, 2, 1.1))
    lam, mu, phi, psi, c = _mat(m)
    ctx = sm.context()
    u = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=8.0)
    eps = O.multiply(B, ix["dof_indexes"], u)
    ws_o, wd_o, ep_o, al_o, mt_o, br_o = O.compute_sigma(eps, ix["weight"], lam, mu, phi, psi, c, dep=True)
    ws_g = torch.zeros(B.rows, dtype=torch.float64, device=gpu)
    wd_g = torch.zeros(3 * B.rows, dtype=torch.float64, device=gpu)
    ctx.compute_sigma_dep(_t(eps, gpu), ws_g, wd_g)
    ctx.synchronize()
    assert np.array_equal(_np(ws_g), ws_o)
    assert np.array_equal(_np(wd_g), wd_o)


def test_random_states_all_branches(gpu):
    """Random strain / plastic-state pairs covering every return branch: flags,
    stresses and plastic strains bit-exact."""
    import torch
    m, sm, B, ix = _setup((4, 4, 2, 1.0))
    lam, mu, phi, psi, c = _mat(m)
    ctx = sm.context()
    npts = B.rows // 3
    rng = np.random.default_rng(11)
    eps = rng.standard_normal((npts, 3)) * 0.02
    ep = np.zeros((npts, 6))
    ep[:, [0, 1, 3, 5]] = rng.standard_normal((npts, 4)) * 0.005
    ctx.set_state(ep)
    ws_o, ep_o, al_o, mt_o, br_o = O.compute_sigma(eps, ix["weight"], lam, mu, phi, psi, c, plastic_strain6=ep,
                                                   update_mem=True)
    assert set(np.unique(br_o)) >= {0, 1, 4}, np.unique(br_o)
    ws_g = torch.zeros(B.rows, dtype=torch.float64, device=gpu)
    ctx.set_update_mem(True)
    ctx.compute_sigma(_t(eps.reshape(-1), gpu), ws_g)
    ctx.set_update_mem(False)
    ctx.synchronize()
    ep_g, al_g, mt_g, br_g = ctx.get_state()
    assert np.array_equal(br_g, br_o)
    assert np.array_equal(mt_g, mt_o)
    assert np.array_equal(_np(ws_g), ws_o)
    assert np.array_equal(ep_g, ep_o)
