"""The patch-assembled residual (ipf_set_residual_assembly(ctx, patch))
against the colored one on the GPU.

Bars:
  * same residual up to the order of the per-node sums: <= 1e-13 of
    max|f_int| (measured ~1e-16), exact and certified-elastic stress modes,
    Q1 / Q2 / Q3 meshes, the mixed triangle / quadrilateral Gmsh mesh and a
    batch of uneven instances (patches never mix materials);
  * r entirely overwritten (NaN before the call), deterministic (repeated
    calls bit-identical);
  * committing residuals keep the colored path: the committed state is
    bit-identical with either assembly selected; after a commit the patch
    residual sees the new state (its plastic-strain copy is refreshed).
"""
import os

import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import batch as bt
from intpointsfem_amd import wellbore as wb

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _res(ctx, u, neq, dev, commit=False):
    import torch
    r = torch.full((neq,), float("nan"), dtype=torch.float64, device=dev)
    ctx.set_update_mem(commit)
    ctx.residual(u, r)
    ctx.set_update_mem(False)
    ctx.synchronize()
    return r.cpu().numpy()


def _compare(ctx, neq, flin, u1, u2, dev, label):
    # state 1 committed through the colored path, then residuals at u2 both ways
    ctx.set_residual_assembly("colored")
    _res(ctx, u1, neq, dev, commit=True)
    st1 = ctx.get_state()
    out = {}
    for mode in ("exact", "certified-elastic"):
        ctx.set_residual_mode(mode)
        ctx.set_residual_assembly("colored")
        rc = _res(ctx, u2, neq, dev)
        ctx.set_residual_assembly("patch")
        rp = _res(ctx, u2, neq, dev)
        rp2 = _res(ctx, u2, neq, dev)
        scale = np.abs(rc - flin).max()
        err = np.abs(rp - rc).max() / scale
        print(f"{label} {mode}: patch vs colored {err:.2e} of max|f_int|", flush=True)
        assert np.isfinite(rp).all()
        assert np.array_equal(rp, rp2)
        assert err <= 1e-13
        out[mode] = rp
    ctx.set_residual_mode("exact")
    # a commit with the patch assembly selected runs the colored path
    ctx.set_residual_assembly("patch")
    rcommit = _res(ctx, u2, neq, dev, commit=True)
    stp = ctx.get_state()
    ctx.set_residual_assembly("colored")
    ctx.set_state(st1[0], st1[1])
    rcommit_c = _res(ctx, u2, neq, dev, commit=True)
    stc = ctx.get_state()
    assert np.array_equal(rcommit, rcommit_c)
    assert all(np.array_equal(a, b) for a, b in zip(stp, stc))
    # after the commit: the patch residual uses the new state
    rc = _res(ctx, u2, neq, dev)
    ctx.set_residual_assembly("patch")
    rp = _res(ctx, u2, neq, dev)
    ctx.set_residual_assembly("colored")
    assert np.abs(rp - rc).max() <= 1e-13 * np.abs(rc - flin).max()
    return out


@pytest.mark.parametrize("mesh", [(40, 36, 1, 1.1), "40x40-q2", (20, 18, 3, 1.2)],
                         ids=lambda t: t if isinstance(t, str) else f"{t[0]}x{t[1]}-q{t[2]}")
def test_patch_residual_equals_colored(gpu, mesh):
    import torch
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    ctx = sm.context()
    neq = sm.sizes()["neq"]
    flin = sm.export_indexes()["rhs_linear"]
    base = wb.lame_displacement(m.nodes(), wb.normal_stress(10))
    u1 = torch.as_tensor(2.0 * base, device=gpu)
    u2 = torch.as_tensor(3.0 * base, device=gpu)
    ctx.set_residual_assembly("patch")
    info = ctx.patch_info()
    print(f"{mesh}: {info}")
    assert info["built"] == 1 and info["patches"] > 0
    _compare(ctx, neq, flin, u1, u2, gpu, str(mesh))


def test_patch_residual_mixed_mesh_and_batch(gpu):
    import torch
    mix = ipf.CompMesh.from_gmsh(os.path.join(HERE, "golden", "gmsh", "wellbore_mix.msh"), 2)
    mix.set_material(5000, 0.2, 2.0, 0.5, 0.5)
    mix.set_normal_stress(-20.0, 0.0)
    sm = ipf.StructMatrix(mix, device=0)
    ctx = sm.context()
    neq = sm.sizes()["neq"]
    rng = np.random.default_rng(1)
    xy = mix.nodes()
    r = np.hypot(xy[:, 0], xy[:, 1])
    u = np.zeros(neq)
    u[0::2] = -2e-3 * xy[:, 0] / r ** 2
    u[1::2] = -2e-3 * xy[:, 1] / r ** 2
    u1 = torch.as_tensor(u, device=gpu)
    u2 = torch.as_tensor(1.7 * u + 1e-6 * rng.standard_normal(neq), device=gpu)
    ctx.set_residual_assembly("patch")
    _compare(ctx, neq, sm.export_indexes()["rhs_linear"], u1, u2, gpu, "wellbore_mix")
    table = [dict(N=9, p=1, phi_deg=25.0, cohesion=3.0, E=3000.0, nu=0.2, pw_final=-20.0),
             dict(N=5, p=3, phi_deg=35.0, cohesion=2.0, E=15000.0, nu=0.28, pw_final=-28.0),
             dict(N=11, p=2, phi_deg=20.0, cohesion=1.5, E=8000.0, nu=0.16, pw_final=-16.0),
             dict(N=7, p=2, phi_deg=38.0, cohesion=6.0, E=2500.0, nu=0.25, pw_final=-29.0)]
    meshes = [bt.instance_mesh(i) for i in table]
    B = bt.Batch(meshes, device=0)
    us = [wb.lame_displacement(mm.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, i["pw_final"], wb.SIGMA0),
                               E_=i["E"], nu=i["nu"]) for i, mm in zip(table, meshes)]
    u1 = torch.as_tensor(np.concatenate([2 * x for x in us]), device=gpu)
    u2 = torch.as_tensor(np.concatenate([4 * x for x in us]), device=gpu)
    B.ctx.set_residual_assembly("patch")
    _compare(B.ctx, B.neq, B.rhs_linear, u1, u2, gpu, "batch")


def test_patch_assembly_needs_geometry(gpu):
    """Uploaded blocks (ipf_transfer_data) carry no corners: the patch
    assembly is refused with a clear error, the colored one keeps working."""
    m = wb.make_mesh((6, 5, 2, 1.1))
    sm = ipf.StructMatrix(m, device=-1)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    s = sm.sizes()
    ctx = ipf.DeviceContext(device=0)
    ctx.transfer_data(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                      b["storage"], ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"])
    with pytest.raises(ipf.IPFError, match="geometric setup"):
        ctx.set_residual_assembly("patch")
    assert ctx.residual_assembly() == "colored"


@pytest.mark.parametrize("waves", [2, 8])
def test_patch_waves_give_the_same_residual(gpu, waves):
    """Patches of 2 and 8 wavefront tiles (4 is the default, covered above):
    the same residual as the colored assembly to the same bar, both stress
    modes, Q2 and Q3."""
    import torch
    for mesh in ("40x40-q2", (20, 18, 3, 1.2)):
        m = wb.make_mesh(mesh)
        sm = ipf.StructMatrix(m, device=0)
        ctx = sm.context()
        neq = sm.sizes()["neq"]
        flin = sm.export_indexes()["rhs_linear"]
        base = wb.lame_displacement(m.nodes(), wb.normal_stress(10))
        ctx.set_patch_waves(waves)
        assert ctx.patch_waves() == waves
        _compare(ctx, neq, flin, torch.as_tensor(2.0 * base, device=gpu), torch.as_tensor(3.0 * base, device=gpu),
                 gpu, f"{mesh} waves={waves}")
    with pytest.raises(ipf.IPFError, match="2, 4 or 8"):
        ctx.set_patch_waves(6)


def test_autotune_assembly_keeps_the_fastest(gpu):
    """ipf_autotune_assembly times colored and patch-2/4/8, keeps the
    fastest, leaves the committed state alone, and the residual of the
    chosen layout matches the colored one (1e-13 of max|f_int|).  Without the
    geometric setup only the colored assembly is timed."""
    import torch
    m = wb.make_mesh("40x40-q2")
    sm = ipf.StructMatrix(m, device=0)
    ctx = sm.context()
    neq = sm.sizes()["neq"]
    flin = sm.export_indexes()["rhs_linear"]
    base = wb.lame_displacement(m.nodes(), wb.normal_stress(10))
    u1 = torch.as_tensor(2.0 * base, device=gpu)
    u2 = torch.as_tensor(3.0 * base, device=gpu)
    _res(ctx, u1, neq, gpu, commit=True)
    st = ctx.get_state()
    rc = _res(ctx, u2, neq, gpu)
    out = ctx.autotune_assembly(u2, reps=5)
    print(out)
    t = out["times_ms"]
    assert set(t) == {"colored", "patch-2", "patch-4", "patch-8"}
    assert out["selected"] == min(t, key=t.get)
    assert all(np.array_equal(a, b) for a, b in zip(st, ctx.get_state()))
    r = _res(ctx, u2, neq, gpu)
    assert np.abs(r - rc).max() <= 1e-13 * np.abs(rc - flin).max()
    # uploaded blocks (no geometry): only the colored assembly is timed
    h = ipf.StructMatrix(m, device=-1)
    b = h.export_blocks()
    ix = h.export_indexes()
    sz = h.sizes()
    ctx2 = ipf.DeviceContext(device=0)
    ctx2.transfer_data(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                       b["storage"], ix["dof_indexes"], ix["color_indexes"], sz["ncolor"], sz["neq"], ix["weight"])
    mat = m.material()
    ctx2.set_material(mat["lambda"], mat["mu"], mat["phi"], mat["psi"], mat["cohesion"])
    ctx2.set_rhs_linear(ix["rhs_linear"])
    out2 = ctx2.autotune_assembly(u2, reps=2)
    assert set(out2["times_ms"]) == {"colored"} and out2["selected"] == "colored"


def test_patch_residual_timed_reports_the_state_refresh(gpu):
    """ipf_residual_timed over the patch layout: launch 0 (the refresh of the
    patch copy of the plastic strain) takes a positive time after a commit
    and reads exactly 0 once the copy is current; the fix-up is the last
    launch (0 points); the residual equals the untimed one bit for bit."""
    import torch
    m = wb.make_mesh("40x40-q2")
    sm = ipf.StructMatrix(m, device=0)
    ctx = sm.context()
    neq = sm.sizes()["neq"]
    base = wb.lame_displacement(m.nodes(), wb.normal_stress(10))
    u1 = torch.as_tensor(2.0 * base, device=gpu)
    u2 = torch.as_tensor(3.0 * base, device=gpu)
    ctx.set_residual_assembly("patch")
    _res(ctx, u1, neq, gpu, commit=True)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ms1, pts1 = ctx.residual_timed(u2, r)
    ms2, pts2 = ctx.residual_timed(u2, r)
    ctx.synchronize()
    rt = r.cpu().numpy()
    print(f"after a commit: {ms1}; current: {ms2}")
    assert ms1[0] > 0.0 and ms2[0] == 0.0
    assert pts2[0] == 0 and pts2[-1] == 0 and pts2.sum() == sm.sizes()["npts"]
    assert np.array_equal(rt, _res(ctx, u2, neq, gpu))
