"""B storage formats (ipf_set_b_format) on the GPU.

Bars:
  * compact FP64: residual, committed flags and plastic strain equal to the
    dense FP64 path (==, i.e. bit-identical up to the sign of a zero), and
    the same K v;
  * FP32 formats (B products in FP32 with FP32 accumulation): residual within
    5e-5 of the largest internal force entry of the FP64 oracle residual
    (measured 8e-6 on the 40 x 40 Q2 wellbore: B.u in FP32 cancels nodal
    displacements ~r/h ~ 25 times larger than the strain near the wall, and
    the assembled forces cancel element contributions once more), branch
    flags differing only at points within the FP32 strain error of the yield
    surface (< 0.5 % of the points).  The FP32 residual is a speed option; the
    converged answer comes from the mixed Newton below;
  * mixed Newton (FP32 CG matvec inside, FP64 residual outside = iterative
    refinement): converges to the FP64 Newton solution within 1e-9 relative
    (both stop at ||r|| <= 1e-12 ||r0|| of the same FP64 residual).
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(mesh):
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    return m, sm, sm.context(), B, ix


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def _residual(ctx, u, neq, dev, commit=False):
    import torch
    r = torch.zeros(neq, dtype=torch.float64, device=dev)
    ctx.set_update_mem(commit)
    ctx.residual(u, r)
    ctx.set_update_mem(False)
    ctx.synchronize()
    return r.cpu().numpy()


@pytest.mark.parametrize("mesh", [(6, 5, 1, 1.2), (40, 40, 2, 1.1), (5, 4, 3, 1.3)],
                         ids=lambda t: f"{t[0]}x{t[1]}-q{t[2]}")
def test_compact_fp64_equals_dense(gpu, mesh):
    m, sm, ctx, B, ix = _setup(mesh)
    neq = m.info()["neq"]
    npts = sm.sizes()["npts"]
    u = _t(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=9.0), gpu)
    ctx.set_b_format("dense-fp64")  # compact FP64 is the default for these blocks
    r_d = _residual(ctx, u, neq, gpu, commit=True)
    ep_d, al_d, mt_d, br_d = ctx.get_state()
    ctx.set_state(np.zeros(6 * npts))
    ctx.set_b_format("compact-fp64")
    r_c = _residual(ctx, u, neq, gpu, commit=True)
    ep_c, al_c, mt_c, br_c = ctx.get_state()
    assert (r_c == r_d).all()
    assert np.array_equal(br_c, br_d) and np.array_equal(mt_c, mt_d)
    assert (ep_c == ep_d).all() and (al_c == al_d).all()
    assert (mt_d != 0).any()


def test_compact_fp64_tangent_equals_dense(gpu):
    import torch
    m, sm, ctx, B, ix = _setup((12, 10, 2, 1.1))
    neq = m.info()["neq"]
    u = _t(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0), gpu)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual_tangent(u, r)
    v = _t(np.random.default_rng(2).standard_normal(neq), gpu)
    y_d = torch.zeros(neq, dtype=torch.float64, device=gpu)
    y_c = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.set_b_format("dense-fp64")
    ctx.tangent_apply(v, y_d)
    ctx.set_b_format("compact-fp64")
    ctx.tangent_apply(v, y_c)
    ctx.synchronize()
    assert (y_c.cpu().numpy() == y_d.cpu().numpy()).all()


@pytest.mark.parametrize("fmt", ["dense-fp32", "compact-fp32"])
def test_fp32_formats_close_to_fp64(gpu, fmt):
    m, sm, ctx, B, ix = _setup("40x40-q2")
    neq = m.info()["neq"]
    s = sm.sizes()
    mat = m.material()
    u_h = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=3.0)
    r_o, ep_o, al_o, mt_o, br_o = O.residual(B, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], neq,
                                             ix["weight"], mat["lambda"], mat["mu"], mat["phi"], mat["psi"],
                                             mat["cohesion"], u_h, f_lin=ix["rhs_linear"], update_mem=True)
    ctx.set_b_format(fmt)
    r_g = _residual(ctx, _t(u_h, gpu), neq, gpu, commit=True)
    _, _, mt_g, br_g = ctx.get_state()
    fint = np.abs(r_o - ix["rhs_linear"]).max()
    err = np.abs(r_g - r_o).max() / fint
    mism = int((br_g != br_o).sum())
    print(f"{fmt}: residual error {err:.2e} of max |f_int|, {mism} branch flags differ of {len(br_o)}")
    assert err <= 5e-5
    assert mism <= 0.005 * len(br_o)


def test_mixed_newton_converges_to_fp64_solution(gpu):
    """Newton with FP32 tangent products (compact FP32 B in the CG matvecs)
    and FP64 residuals -- an iterative refinement -- reaches the FP64 Newton
    solution: both stop at ||r|| <= 1e-12 ||r0|| of the same FP64 residual."""
    import torch
    sols, its = {}, {}
    for tangent in ("dense-fp64", "compact-fp32"):
        m = wb.make_mesh((20, 20, 2, 1.1 ** 2))
        sm = ipf.StructMatrix(m, device=0)
        ix = sm.export_indexes()
        ctx = sm.context()
        ctx.set_dirichlet_mask(m.dirichlet_mask())
        ctx.set_b_format("dense-fp64", tangent=tangent)
        neq = m.info()["neq"]
        f_unit = ix["rhs_linear"] / wb.normal_stress(wb.NSTEPS)
        u = torch.zeros(neq, dtype=torch.float64, device=gpu)
        its[tangent] = []
        for k in (4, 7, 10):
            ctx.set_rhs_linear(f_unit * wb.normal_stress(k))
            st = ctx.newton(u, rtol=1e-12, maxit=80, cg_rtol=1e-6, cg_maxit=20000)
            assert st["converged"] == 1.0, (tangent, k, st)
            its[tangent].append(int(st["iterations"]))
        ctx.synchronize()
        sols[tangent] = u.cpu().numpy()
    ref = sols["dense-fp64"]
    err = np.abs(sols["compact-fp32"] - ref).max() / np.abs(ref).max()
    print(f"newton iterations {its}; mixed vs FP64 solution {err:.2e}")
    assert err <= 1e-9


def test_format_errors(gpu):
    m, sm, ctx, B, ix = _setup((4, 3, 2, 1.2))
    with pytest.raises(ipf.IPFError, match="unknown format"):
        ctx.set_b_format(7)
    with pytest.raises(ipf.IPFError, match="unknown format"):
        ctx.set_b_format(0, tangent=-1)


def test_compact_rejected_for_general_blocks(gpu):
    """A block entry outside the 2D elasticity pattern (here B[xx, uy] != 0)
    rules out the compact formats; the dense FP32 one still works."""
    import torch
    m = wb.make_mesh((4, 3, 2, 1.2))
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    s = sm.sizes()
    storage = b["storage"].copy()
    # block 0 is row-major rows x cols: entry (row 0 = xx of point 0, col 1 = uy of node 0)
    storage[b["matrix_position"][0] + 1] = 0.25
    ctx = ipf.DeviceContext(device=0)
    ctx.transfer_data(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], storage,
                      ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"])
    assert ctx.b_format() == ("dense-fp64", "dense-fp64")
    with pytest.raises(ipf.IPFError, match="compact B"):
        ctx.set_b_format("compact-fp64")
    with pytest.raises(ipf.IPFError, match="compact B"):
        ctx.set_b_format("dense-fp64", tangent="compact-fp32")
    ctx.set_b_format("dense-fp32")
    u = torch.zeros(s["neq"], dtype=torch.float64, device=gpu)
    eps = torch.zeros(s["rows"], dtype=torch.float64, device=gpu)
    ctx.multiply(u, eps)
    ctx.synchronize()


def test_default_formats(gpu):
    """Device setup from geometry: the geometric form; host blocks in the 2D
    pattern: compact FP64."""
    m, sm, ctx, B, ix = _setup((4, 3, 2, 1.2))
    assert ctx.b_format() == ("geometry-fp64", "geometry-fp64")
    ctx.set_b_format("dense-fp64", tangent="compact-fp32")
    assert ctx.b_format() == ("dense-fp64", "compact-fp32")
    b = sm.export_blocks()
    s = sm.sizes()
    ch = ipf.DeviceContext(device=0)
    ch.transfer_data(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                     b["storage"], ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"])
    assert ch.b_format() == ("compact-fp64", "compact-fp64")


def test_autotune_selects_the_fastest_format(gpu):
    """ipf_autotune times every exact FP64 form (dense, compact, geometric)
    for the residual and the tangent product, selects the fastest of each,
    and leaves the committed state alone; the residual afterwards is the same
    bits as before.  With allow_fp32 the FP32 forms compete for the tangent
    product only."""
    import torch
    m, sm, ctx, B, ix = _setup((20, 20, 2, 1.1))
    s = sm.sizes()
    neq = s["neq"]
    u = _t(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=4.0), gpu)
    r0 = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual(u, r0)
    st0 = ctx.get_state()
    rt = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual_tangent(u, rt)
    t = ctx.autotune(u, reps=5)
    assert set(t["residual"]) == {"dense-fp64", "compact-fp64", "geometry-fp64"}
    assert set(t["tangent"]) == {"dense-fp64", "compact-fp64", "geometry-fp64"}
    r_sel, t_sel = t["selected"]
    assert t["residual"][r_sel] == min(t["residual"].values())
    assert t["tangent"][t_sel] == min(t["tangent"].values())
    r1 = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual(u, r1)
    ctx.synchronize()
    assert (r1.cpu().numpy() == r0.cpu().numpy()).all()
    assert all(np.array_equal(a, b) for a, b in zip(ctx.get_state(), st0))
    t2 = ctx.autotune(u, reps=5, allow_fp32=True)
    assert "dense-fp32" not in t2["residual"] and "compact-fp32" not in t2["residual"]
    assert {"dense-fp32", "compact-fp32"} <= set(t2["tangent"])
    assert t2["tangent"][t2["selected"][1]] == min(t2["tangent"].values())
    print(f"autotune: residual {t['residual']} -> {r_sel}; tangent {t2['tangent']} -> {t2['selected'][1]}")


def test_geometry_fp32(gpu):
    """geometry-fp32: B recomputed from the corners in FP64 and rounded to
    FP32 is the compact FP32 B, so residuals (FP32 products, FP64 return map)
    equal compact FP32's bit for bit; its tangent product (FP32 w Dep too) is
    the FP64 product to FP32 accuracy (1e-5 of max |K v|); a Newton with it in
    the CG reaches the FP64 Newton's solution (iterative refinement)."""
    import torch
    m, sm, ctx, B, ix = _setup((20, 20, 2, 1.1))
    neq = m.info()["neq"]
    u = _t(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=4.0), gpu)
    ctx.set_b_format("compact-fp32")
    r_c = _residual(ctx, u, neq, gpu)
    ctx.set_b_format("geometry-fp32")
    r_g = _residual(ctx, u, neq, gpu)
    assert np.array_equal(r_c, r_g)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual_tangent(u, r)
    v = _t(np.random.default_rng(0).standard_normal(neq), gpu)
    y = {}
    for fmt in ("geometry-fp64", "geometry-fp32"):
        ctx.set_b_format("geometry-fp64", tangent=fmt)
        yy = torch.full((neq,), float("nan"), dtype=torch.float64, device=gpu)
        ctx.tangent_apply(v, yy)
        ctx.synchronize()
        y[fmt] = yy.cpu().numpy()
    err = np.abs(y["geometry-fp32"] - y["geometry-fp64"]).max() / np.abs(y["geometry-fp64"]).max()
    print(f"geometry-fp32 tangent product vs FP64: {err:.2e}")
    assert 0 < err <= 1e-5
    sols = {}
    for tangent in ("geometry-fp64", "geometry-fp32"):
        m2 = wb.make_mesh((20, 20, 2, 1.1 ** 2))
        sm2 = ipf.StructMatrix(m2, device=0)
        c2 = sm2.context()
        c2.set_dirichlet_mask(m2.dirichlet_mask())
        c2.set_b_format("geometry-fp64", tangent=tangent)
        u2 = torch.zeros(m2.info()["neq"], dtype=torch.float64, device=gpu)
        st = c2.newton(u2, rtol=1e-12, maxit=80, cg_rtol=1e-6, cg_maxit=20000)
        assert st["converged"] == 1.0, (tangent, st)
        sols[tangent] = u2.cpu().numpy()
    ref = sols["geometry-fp64"]
    e2 = np.abs(sols["geometry-fp32"] - ref).max() / np.abs(ref).max()
    print(f"newton with geometry-fp32 tangent vs FP64: {e2:.2e}")
    assert e2 <= 1e-9
