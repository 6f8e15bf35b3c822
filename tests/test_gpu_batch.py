"""Batched instances of uneven size, order and material in one context.

Bar: every instance's slice of the batched residual, and its committed
branch flags, mtype and plastic strain, are bit-identical to the same
instance computed alone (each residual entry still gets its element
contributions in color order, one per launch); the single-instance path is
itself pinned against the oracle by the parity tests.  FP32 and dense
formats hold the same property.  The batched Newton (one global system,
block diagonal) reaches every instance's own Newton solution.
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import batch
from intpointsfem_amd import wellbore as wb

pytestmark = pytest.mark.gpu

TABLE = [dict(N=9, p=1, phi_deg=25.0, cohesion=3.0, E=3000.0, nu=0.2, pw_final=-20.0),
         dict(N=5, p=3, phi_deg=35.0, cohesion=2.0, E=15000.0, nu=0.28, pw_final=-28.0),
         dict(N=11, p=2, phi_deg=20.0, cohesion=1.5, E=8000.0, nu=0.16, pw_final=-16.0),
         dict(N=7, p=2, phi_deg=38.0, cohesion=6.0, E=2500.0, nu=0.25, pw_final=-29.0),
         dict(N=8, p=1, phi_deg=30.0, cohesion=1.0, E=19000.0, nu=0.3, pw_final=-15.0)]


def _u(inst, m, scale):
    return wb.lame_displacement(m.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, inst["pw_final"], wb.SIGMA0),
                                scale=scale, E_=inst["E"], nu=inst["nu"])


@pytest.mark.parametrize("fmt", ["compact-fp64", "dense-fp64", "compact-fp32"])
def test_batch_equals_instances_alone(gpu, fmt):
    import torch
    meshes = [batch.instance_mesh(i) for i in TABLE]
    B = batch.Batch(meshes, device=0)
    us = [_u(inst, m, 4.0) for inst, m in zip(TABLE, meshes)]
    u = torch.as_tensor(np.concatenate(us), dtype=torch.float64, device=gpu)
    r = torch.zeros(B.neq, dtype=torch.float64, device=gpu)
    B.ctx.set_b_format(fmt)
    B.ctx.set_update_mem(True)
    B.ctx.residual(u, r)
    B.ctx.set_update_mem(False)
    B.ctx.synchronize()
    rb = r.cpu().numpy()
    ep_b, al_b, mt_b, br_b = B.ctx.get_state()
    assert (mt_b != 0).any() and (mt_b == 0).any()
    for i, (inst, m) in enumerate(zip(TABLE, meshes)):
        sm = ipf.StructMatrix(m, device=0)
        ctx = sm.context()
        ctx.set_b_format(fmt)
        ui = torch.as_tensor(us[i], dtype=torch.float64, device=gpu)
        ri = torch.zeros(len(us[i]), dtype=torch.float64, device=gpu)
        ctx.set_update_mem(True)
        ctx.residual(ui, ri)
        ctx.set_update_mem(False)
        ctx.synchronize()
        e0, e1 = B.eq_offsets[i], B.eq_offsets[i + 1]
        p0, p1 = B.pt_offsets[i], B.pt_offsets[i + 1]
        assert (rb[e0:e1] == ri.cpu().numpy()).all(), i
        ep, al, mt, br = ctx.get_state()
        assert np.array_equal(mt_b[p0:p1], mt) and np.array_equal(br_b[p0:p1], br), i
        assert (ep_b[6 * p0:6 * p1] == ep).all() and (al_b[p0:p1] == al).all(), i


# instances strong enough to carry their final load (perfect plasticity has a
# limit load: the weak entries of TABLE have no equilibrium at full load)
NEWTON_TABLE = [dict(N=9, p=1, phi_deg=30.0, cohesion=10.0, E=5000.0, nu=0.2, pw_final=-19.5),
                dict(N=5, p=3, phi_deg=35.0, cohesion=12.0, E=15000.0, nu=0.28, pw_final=-25.0),
                dict(N=8, p=2, phi_deg=28.0, cohesion=8.0, E=8000.0, nu=0.16, pw_final=-22.0)]


def test_batch_newton_reaches_each_instance_solution(gpu):
    """10 load steps (the reference driver's ramp) of the batched Newton --
    one global system, block diagonal -- against each instance's own Newton:
    with the global criterion (ipf_newton) within 1e-7 of max |u_i|, and with
    per-instance convergence (ipf_newton_segmented: every instance to its own
    ||r_i|| <= 1e-10 ||r0_i|| at every step) within 1e-12 (measured 4e-16 to
    9e-16 on MI355X, against 1e-13 to 2e-12 with the global criterion); and
    per-instance convergence with the loose CG tolerance 0.1 (every step
    converges; within 1e-7 of the instance alone, the bar of the global
    criterion)."""
    import torch
    table = NEWTON_TABLE
    meshes = [batch.instance_mesh(i) for i in table]
    B = batch.Batch(meshes, device=0)
    B.ctx.set_dirichlet_mask(np.concatenate([m.dirichlet_mask() for m in meshes]))
    u = torch.zeros(B.neq, dtype=torch.float64, device=gpu)
    for k in range(1, wb.NSTEPS + 1):
        B.ctx.set_rhs_linear(B.rhs_linear * (k / wb.NSTEPS))
        st = B.ctx.newton(u, rtol=1e-10, maxit=80, cg_rtol=1e-12, cg_maxit=40000)
        assert st["converged"] == 1.0, (k, st)
    _, _, mt, _ = B.ctx.get_state()
    assert (mt != 0).any()
    ub = u.cpu().numpy()
    # per-instance convergence from the same start in a fresh batch
    Bs = batch.Batch(meshes, device=0)
    Bs.ctx.set_dirichlet_mask(np.concatenate([m.dirichlet_mask() for m in meshes]))
    us = torch.zeros(Bs.neq, dtype=torch.float64, device=gpu)
    for k in range(1, wb.NSTEPS + 1):
        Bs.ctx.set_rhs_linear(Bs.rhs_linear * (k / wb.NSTEPS))
        st = Bs.newton(us, rtol=1e-10, maxit=80, cg_rtol=1e-12, cg_maxit=40000)
        assert st["converged"] == 1.0, (k, st)
        assert (st["seg_rel"] <= 1e-10).all() and (st["seg_iters"] >= 0).all(), (k, st)
        print(f"segmented step {k}: newton iterations per instance {st['seg_iters'].tolist()}", flush=True)
    useg = us.cpu().numpy()
    # the same with the loose inner tolerance of the bench's solves (0.1): the
    # segments' right-hand sides are scaled to unit norm inside, so each one
    # gets the CG accuracy it would get alone and every step converges
    Bl = batch.Batch(meshes, device=0)
    Bl.ctx.set_dirichlet_mask(np.concatenate([m.dirichlet_mask() for m in meshes]))
    ul = torch.zeros(Bl.neq, dtype=torch.float64, device=gpu)
    for k in range(1, wb.NSTEPS + 1):
        Bl.ctx.set_rhs_linear(Bl.rhs_linear * (k / wb.NSTEPS))
        st = Bl.newton(ul, rtol=1e-10, maxit=80, cg_rtol=1e-1, cg_maxit=40000)
        print(f"loose-CG segmented step {k}: newton iterations per instance {st['seg_iters'].tolist()}", flush=True)
        assert st["converged"] == 1.0 and (st["seg_iters"] >= 0).all(), (k, st)
    uloose = ul.cpu().numpy()
    for i, m in enumerate(meshes):
        sm = ipf.StructMatrix(m, device=0)
        ctx = sm.context()
        ctx.set_dirichlet_mask(m.dirichlet_mask())
        f = sm.export_indexes()["rhs_linear"]
        ui = torch.zeros(sm.sizes()["neq"], dtype=torch.float64, device=gpu)
        for k in range(1, wb.NSTEPS + 1):
            ctx.set_rhs_linear(f * (k / wb.NSTEPS))
            si = ctx.newton(ui, rtol=1e-10, maxit=80, cg_rtol=1e-12, cg_maxit=40000)
            assert si["converged"] == 1.0, (i, k, si)
        ref = ui.cpu().numpy()
        err = np.abs(ub[B.eq_offsets[i]:B.eq_offsets[i + 1]] - ref).max() / np.abs(ref).max()
        err_s = np.abs(useg[Bs.eq_offsets[i]:Bs.eq_offsets[i + 1]] - ref).max() / np.abs(ref).max()
        err_l = np.abs(uloose[Bl.eq_offsets[i]:Bl.eq_offsets[i + 1]] - ref).max() / np.abs(ref).max()
        print(f"instance {i}: |u_batch - u_alone| / |u| = {err:.2e} (global criterion), {err_s:.2e} (per instance), "
              f"{err_l:.2e} (per instance, CG tolerance 0.1)")
        assert err <= 1e-7
        assert err_s <= 1e-12
        assert err_l <= 1e-7


def test_material_errors(gpu):
    import torch
    meshes = [batch.instance_mesh(i) for i in TABLE[:2]]
    h = batch.concat_instances(meshes)
    ctx = ipf.DeviceContext(device=0)
    ctx.set_element_materials(h["element_material"])
    ctx.transfer_data(h["row_sizes"], h["col_sizes"], h["matrix_position"], h["row_first"], h["col_first"],
                      h["storage"], h["dof_indexes"], h["color_indexes"], h["ncolor"], h["neq"], h["weight"])
    ctx.set_material_table(h["material_table"][:1])  # material 1 missing
    z = torch.zeros(h["neq"], dtype=torch.float64, device=gpu)
    with pytest.raises(ipf.IPFError, match="material 1 not set"):
        ctx.residual(z, z)
    with pytest.raises(ipf.IPFError, match="material id"):
        ctx.set_element_materials(-h["element_material"] - 1)
    with pytest.raises(ipf.IPFError, match="bad material"):
        ctx.set_material_table([[1.0, -1.0, 0.5, 0.5, 1.0]])
