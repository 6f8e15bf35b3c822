"""GPU tangent, CG and Newton against the CPU oracle (oracle/, oracle/newton.py).

Bars:
  * K v (matrix-free tangent) vs the oracle's assembled K: 1e-12 relative to
    max |K v| (different summation order only);
  * CG: ||K x - b|| <= 1e-9 ||b|| measured with the oracle's K;
  * 10-step elastoplastic Newton (40 x 40 Q2 wellbore): per-step displacements
    within 1e-10 relative (both sides converge to ||r|| <= 1e-12 ||r0||; measured
    ~1e-15), committed branch flags
    bit-identical;
  * elastic limit (c -> inf): the FE displacement matches the thick-cylinder
    closed form to 5e-4 relative (the straight-sided cells approximate the
    arcs with 40 chords; measured 2.6e-4).
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb
from oracle import oracle as O
from oracle.newton import OracleNewton, assemble_tangent

pytestmark = pytest.mark.gpu


def _setup(mesh, cohesion=None):
    m = wb.make_mesh(mesh)
    if cohesion is not None:
        m.set_material(wb.E, wb.NU, cohesion, wb.PHI, wb.PSI)
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    ctx = sm.context()
    mask = m.dirichlet_mask()
    ctx.set_dirichlet_mask(mask)
    return m, sm, ctx, B, ix, mask


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)


def test_tangent_apply_matches_assembled_oracle(gpu):
    import torch
    m, sm, ctx, B, ix, mask = _setup((12, 10, 2, 1.1))
    neq = m.info()["neq"]
    u = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual_tangent(_t(u, gpu), r)
    mat = m.material()
    strain = O.multiply(B, ix["dof_indexes"], u)
    # update_mem=True only fills the returned (local) state arrays, including mtype
    ws, wd, ep, al, mt, br = O.compute_sigma(strain, ix["weight"], mat["lambda"], mat["mu"], mat["phi"], mat["psi"],
                                             mat["cohesion"], dep=True, update_mem=True)
    assert (mt != 0).any()
    K = assemble_tangent(B, ix["dof_indexes"], wd, neq)
    rng = np.random.default_rng(0)
    v = rng.standard_normal(neq)
    y = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.tangent_apply(_t(v, gpu), y)
    ctx.synchronize()
    ref = K @ v
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    # the residual computed alongside the tangent equals the plain residual
    r2 = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual(_t(u, gpu), r2)
    ctx.synchronize()
    assert np.abs(r.cpu().numpy() - r2.cpu().numpy()).max() <= 1e-12 * np.abs(ix["rhs_linear"]).max()


def test_cg_solves_tangent_system(gpu):
    import torch
    m, sm, ctx, B, ix, mask = _setup((12, 10, 2, 1.1))
    neq = m.info()["neq"]
    u = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual_tangent(_t(u, gpu), r)
    mat = m.material()
    strain = O.multiply(B, ix["dof_indexes"], u)
    _, wd, *_ = O.compute_sigma(strain, ix["weight"], mat["lambda"], mat["mu"], mat["phi"], mat["psi"],
                                mat["cohesion"], dep=True)
    K = assemble_tangent(B, ix["dof_indexes"], wd, neq)
    free = ~mask
    rng = np.random.default_rng(1)
    b = rng.standard_normal(neq)
    x = torch.zeros(neq, dtype=torch.float64, device=gpu)
    its, rel = ctx.cg_solve(_t(b, gpu), x, rtol=1e-11, maxit=20000)
    xs = x.cpu().numpy()
    assert (xs[mask] == 0).all()
    res = (K[free][:, free] @ xs[free]) - b[free]
    assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(b[free]), (its, rel)


def test_newton_load_steps_match_oracle(gpu):
    """Config 1 (40 x 40 Q2, ratio 1.1): 10 load steps on the GPU (Newton +
    block-Jacobi PCG) and on the CPU oracle (Newton + direct solve)."""
    import torch
    m, sm, ctx, B, ix, mask = _setup("40x40-q2")
    s = sm.sizes()
    neq = s["neq"]
    orc = OracleNewton(B, ix, neq, s["ncolor"], m.material(), mask, nthreads=16)
    f_unit = ix["rhs_linear"] / wb.normal_stress(wb.NSTEPS)
    u = torch.zeros(neq, dtype=torch.float64, device=gpu)
    worst = 0.0
    for k in range(1, wb.NSTEPS + 1):
        f = f_unit * wb.normal_stress(k)
        ctx.set_rhs_linear(f)
        st = ctx.newton(u, rtol=1e-12, maxit=40, cg_rtol=1e-13, cg_maxit=40000)
        so = orc.load_step(f, rtol=1e-12, maxit=40)
        assert st["converged"] == 1.0, (k, st)
        assert so["converged"], (k, so)
        ug = u.cpu().numpy()
        rel = np.abs(ug - orc.u).max() / np.abs(orc.u).max()
        worst = max(worst, rel)
        ep, al, mt, br = ctx.get_state()
        print(f"step {k}: newton {int(st['iterations'])} (oracle {so['iterations']}), cg {int(st['cg_iterations'])}, "
              f"|du|/|u| {rel:.2e}, plastic {int((mt != 0).sum())}, flag mismatches {int((br != so['branch']).sum())}",
              flush=True)
        assert rel <= 1e-10, (k, rel)
        assert np.array_equal(br, so["branch"]), k
    assert (mt != 0).sum() > 1000


def test_elastic_limit_matches_thick_cylinder(gpu):
    """c = 1e10: one load step is linear elastic; compare the radial
    displacement with the Lame closed form."""
    import torch
    m, sm, ctx, B, ix, mask = _setup("40x40-q2", cohesion=1e10)
    neq = m.info()["neq"]
    u = torch.zeros(neq, dtype=torch.float64, device=gpu)
    st = ctx.newton(u, rtol=1e-12, maxit=10, cg_rtol=1e-13, cg_maxit=40000)
    assert st["converged"] == 1.0 and st["iterations"] <= 2
    xy = m.nodes()
    ref = wb.lame_displacement(xy, wb.normal_stress(wb.NSTEPS))
    ug = u.cpu().numpy()
    r = np.hypot(xy[:, 0], xy[:, 1])
    ur = (ug[0::2] * xy[:, 0] + ug[1::2] * xy[:, 1]) / r
    ur_ref = (ref[0::2] * xy[:, 0] + ref[1::2] * xy[:, 1]) / r
    err = np.abs(ur - ur_ref).max() / np.abs(ur_ref).max()
    print(f"elastic limit: newton {int(st['iterations'])}, cg {int(st['cg_iterations'])}, max rel u_r error {err:.2e}")
    # the cells are straight-sided: 40 chords stand for each arc, an O(h^2)
    # geometric error of (pi/80)^2 ~ 1.5e-3 on the wall; the field itself is Q2
    assert err <= 5e-4
    _, _, mt, _ = ctx.get_state()
    assert (mt == 0).all()


def test_cg_graph_matches_single_launches(gpu):
    """The Newton iterations replayed from captured HIP graphs give the same
    iterates, bit for bit, as the same kernels launched one by one: 10 load
    steps of config 1 with (a) the Newton step graph (tangent residual,
    residual norm, CG start and first CG block per iteration) and the CG
    block graph, (b) the CG block graph only, (c) everything eager (every
    step compared; the graphs are replayed hundreds of times with eager
    residual / tangent work and host reads between the replays), then a
    B-format change (the graphs are rebuilt) and a further step each with the
    compact FP64 and the FP32 geometric tangent (its FP32 copy of w Dep is
    made inside the step graph)."""
    import time

    import torch
    res = {}
    for mode in ("step+cg", "cg", "eager"):
        m, sm, ctx, B, ix, mask = _setup("40x40-q2")
        ctx.set_cg_graph(mode != "eager")
        ctx.set_newton_graph(mode == "step+cg")
        neq = sm.sizes()["neq"]
        f_unit = ix["rhs_linear"] / wb.normal_stress(wb.NSTEPS)
        u = torch.zeros(neq, dtype=torch.float64, device=gpu)
        steps = []
        t0 = time.time()
        for k in range(1, wb.NSTEPS + 1):
            ctx.set_rhs_linear(f_unit * wb.normal_stress(k))
            st = ctx.newton(u, rtol=1e-12, maxit=40, cg_rtol=1e-13, cg_maxit=40000)
            assert st["converged"] == 1.0, (mode, k, st)
            steps.append((u.cpu().numpy().copy(), int(st["iterations"]), int(st["cg_iterations"]), st["r"]))
        elapsed = time.time() - t0
        for tangent, scale in (("compact-fp64", 1.01), ("geometry-fp32", 1.02)):
            ctx.set_b_format(ctx.b_format()[0], tangent=tangent)
            ctx.set_rhs_linear(f_unit * wb.normal_stress(wb.NSTEPS) * scale)
            st = ctx.newton(u, rtol=1e-10, maxit=40, cg_rtol=1e-8, cg_maxit=40000)
            steps.append((u.cpu().numpy().copy(), int(st["iterations"]), int(st["cg_iterations"]), st["r"]))
        res[mode] = (steps, elapsed, ctx.get_state())
        print(f"{mode}: 10 load steps {elapsed:.2f} s, cg iterations {sum(s[2] for s in steps[:10])}", flush=True)
    for mode in ("step+cg", "cg"):
        for (ug, ig, cg, rg), (ue, ie, ce, re) in zip(res[mode][0], res["eager"][0]):
            assert np.array_equal(ug, ue) and ig == ie and cg == ce and rg == re, mode
        assert all(np.array_equal(a, b) for a, b in zip(res[mode][2], res["eager"][2])), mode


def test_newton_at_a_converged_u_stops_on_the_absolute_floor(gpu):
    """A second ipf_newton on the same load starts at the converged u, where
    r0 is at rounding level and ||r|| <= rtol ||r0|| cannot be met.  With the
    absolute floor (ipf_set_newton_atol; the reference stops on absolute
    norms, main.cpp:235) it returns at once, converged, u unchanged."""
    import torch
    m, sm, ctx, B, ix, mask = _setup((16, 12, 2, 1.1))
    neq = sm.sizes()["neq"]
    f = ix["rhs_linear"] * (wb.normal_stress(3) / wb.normal_stress(wb.NSTEPS))  # load step 3
    ctx.set_rhs_linear(f)
    u = torch.zeros(neq, dtype=torch.float64, device=gpu)
    st = ctx.newton(u, rtol=1e-10, maxit=80, cg_rtol=1e-13, cg_maxit=40000)
    assert st["converged"] == 1.0 and st["iterations"] >= 1, st
    u1 = u.cpu().numpy().copy()
    fnorm = float(np.linalg.norm(f[mask == 0]))
    ctx.set_newton_atol(1e-10 * fnorm)
    st2 = ctx.newton(u, rtol=1e-12, maxit=40, cg_rtol=1e-13, cg_maxit=40000)
    print(f"first call: {int(st['iterations'])} iterations, |r| {st['r']:.3e}; second call: "
          f"{int(st2['iterations'])} iterations, |r| {st2['r']:.3e}, floor {1e-10 * fnorm:.3e}")
    assert st2["converged"] == 1.0 and st2["iterations"] == 0
    np.testing.assert_array_equal(u.cpu().numpy(), u1)
    ctx.set_newton_atol(0.0)
    with pytest.raises(ipf.IPFError):
        ctx.set_newton_atol(-1.0)


def test_newton_restores_update_mem_after_an_error(gpu):
    """ipf_newton turns the caller's UpdateMem off for its Newton residuals;
    an error inside (here: cg_maxit < 1, rejected by ipf_cg_solve) must leave
    it as the caller set it -- a later residual still commits."""
    import torch
    m, sm, ctx, B, ix, mask = _setup((12, 10, 2, 1.1))
    neq = sm.sizes()["neq"]
    ctx.set_rhs_linear(ix["rhs_linear"])
    ctx.set_update_mem(1)
    u = torch.zeros(neq, dtype=torch.float64, device=gpu)
    with pytest.raises(ipf.IPFError, match="maxit"):
        ctx.newton(u, rtol=1e-12, maxit=5, cg_rtol=1e-13, cg_maxit=0)
    ep, _, mt, _ = ctx.get_state()
    assert not ep.any()  # nothing committed by the failed call
    up = _t(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0), gpu)
    r = torch.zeros(neq, dtype=torch.float64, device=gpu)
    ctx.residual(up, r)
    ctx.synchronize()
    ep, _, mt, _ = ctx.get_state()
    assert (mt != 0).any() and ep.any()  # UpdateMem was still on


def test_newton_monitor_reports_every_iteration(gpu):
    """ipf_set_newton_monitor: one call per residual norm, iteration 0 at the
    entry u (rel 1), the last one equal to the returned statistics -- for
    ipf_newton and for ipf_newton_segmented over two equation blocks."""
    import torch
    m, sm, ctx, B, ix, mask = _setup((16, 12, 2, 1.1))
    neq = sm.sizes()["neq"]
    ctx.set_rhs_linear(ix["rhs_linear"] * (wb.normal_stress(3) / wb.normal_stress(wb.NSTEPS)))
    for segmented in (False, True):
        ctx.set_state(*[np.zeros_like(a) for a in ctx.get_state()[:2]])
        calls = []
        ctx.set_newton_monitor(lambda it, rel, rn, cg: calls.append((it, rel, rn, cg)))
        u = torch.zeros(neq, dtype=torch.float64, device=gpu)
        if segmented:
            st = ctx.newton_segmented(u, [0, neq // 2, neq], rtol=1e-10, maxit=80, cg_rtol=1e-13, cg_maxit=40000)
        else:
            st = ctx.newton(u, rtol=1e-10, maxit=80, cg_rtol=1e-13, cg_maxit=40000)
        ctx.set_newton_monitor(None)
        its = [c[0] for c in calls]
        print(f"segmented={segmented}: {len(calls)} reports, final rel {calls[-1][1]:.2e}, stats {st['rel_residual']:.2e}")
        assert its == list(range(int(st["iterations"]) + 1))
        assert calls[0][1] == 1.0 and calls[0][3] == 0
        assert calls[-1][1] == st["rel_residual"] and calls[-1][2] == st.get("rn", st.get("r"))
        assert calls[-1][3] == int(st["cg_iterations"])
        assert all(a[3] <= b[3] for a, b in zip(calls, calls[1:]))
    # off: newton runs without reports
    calls.clear()
    ctx.newton(u, rtol=1e-10, maxit=2, cg_rtol=1e-13, cg_maxit=40000)
    assert not calls


def test_newton_line_search_converges_with_a_loose_inner_tolerance(gpu):
    """ipf_set_newton_line_search: a low-cohesion instance (c = 1.1 MPa,
    phi = 36 deg) over its first 6 load steps with the CG solved to 10 % per
    Newton step.  With the line search every step converges and every
    accepted step lowers ||r|| (or uses all its halvings); the solution
    matches the tight inner solve (CG to 1e-10, full steps) to 1e-5 of
    max |u| (both Newton solves stop at ||r|| <= 1e-8 ||r0||).  (Later load
    steps of this instance need more than 80 such inexact iterations: the
    search keeps them from diverging, not linear.)  With 0 halvings newton()
    takes the full steps, bit for bit."""
    import torch
    from intpointsfem_amd import batch
    t = dict(N=60, p=2, phi_deg=36.0, cohesion=1.1, E=5000.0, nu=0.25, pw_final=-25.0)
    m = batch.instance_mesh(t)
    nsteps = 6
    sols = {}
    for ls, cg in ((8, 1e-1), (0, 1e-10), (-1, 1e-10)):
        sm = ipf.StructMatrix(m, device=0)
        ctx = sm.context()
        ctx.set_dirichlet_mask(m.dirichlet_mask())
        f = sm.export_indexes()["rhs_linear"]
        if ls >= 0:
            ctx.set_newton_line_search(ls)
        u = torch.zeros(sm.sizes()["neq"], dtype=torch.float64, device=gpu)
        halvings = 0
        for k in range(1, nsteps + 1):
            ctx.set_rhs_linear(f * (k / wb.NSTEPS))
            rels = []
            ctx.set_newton_monitor(lambda it, rel, rn, cg_: rels.append(rn))
            st = ctx.newton(u, rtol=1e-8, maxit=80, cg_rtol=cg, cg_maxit=40000)
            ctx.set_newton_monitor(None)
            assert st["converged"] == 1.0, (ls, k, st)
            halvings += int(st["line_search_halvings"])
            if ls > 0 and int(st["line_search_halvings"]) == 0:
                assert all(b < a * (1 + 1e-12) for a, b in zip(rels, rels[1:])), (k, rels)
        _, _, mt, _ = ctx.get_state()
        print(f"line search {ls}: {halvings} halvings, {int((mt != 0).sum())} plastic points")
        sols[ls] = u.cpu().numpy()
    assert np.array_equal(sols[0], sols[-1])  # 0 halvings: the default full steps
    err = np.abs(sols[8] - sols[0]).max() / np.abs(sols[0]).max()
    print(f"|u_ls - u_tight| / max|u| = {err:.2e}")
    assert err <= 1e-5


def test_newton_reference_stop_rule(gpu):
    """The reference's stop test (main.cpp:239, tolerance 1e-7 at :139):
    with ipf_set_newton_reference_stop(1e-7), ipf_newton on load step 1 of
    config 1 stops after N updates at an iterate with ||r|| < 1e-7; a second
    context run with maxit N - 1 reaches iterate N - 1 (the same kernels, the
    same bits) and reports it not converged -- the rule failed there -- and the
    last update u_N - u_(N-1) has norm < 1e-7.  The setter rejects negative
    and non-finite tolerances; 0 restores the relative test."""
    import torch
    tol = 1e-7
    runs = []
    for maxit in (40, None):
        m, sm, ctx, B, ix, mask = _setup("40x40-q2")
        ctx.set_newton_reference_stop(tol)
        ctx.set_rhs_linear(ix["rhs_linear"] / wb.normal_stress(wb.NSTEPS) * wb.normal_stress(1))
        u = torch.zeros(sm.sizes()["neq"], dtype=torch.float64, device=gpu)
        n = maxit if maxit is not None else int(runs[0][1]["iterations"]) - 1
        st = ctx.newton(u, rtol=1e-30, maxit=n, cg_rtol=1e-13, cg_maxit=40000)
        runs.append((u.cpu().numpy(), st, sm, ctx))
        print(f"maxit {n}: {st}", flush=True)
    (ua, sa, sm, ctx), (ub, sb, _, _) = runs
    assert sa["converged"] == 1.0 and sa["iterations"] >= 2 and sa["r"] < tol
    assert sb["converged"] == 0.0 and sb["iterations"] == sa["iterations"] - 1
    assert np.linalg.norm(ua - ub) < tol
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(ipf.IPFError):
            ctx.set_newton_reference_stop(bad)
    # 0 restores the relative test (with its absolute floor): converged at
    # once at the converged u of the first run
    ctx.set_newton_reference_stop(0.0)
    ctx.set_newton_atol(1e-6)
    u = torch.as_tensor(ua, device=gpu)
    st = ctx.newton(u, rtol=1e-12, maxit=5, cg_rtol=1e-13, cg_maxit=40000)
    assert st["converged"] == 1.0 and st["iterations"] == 0, st


def test_newton_thomas_acceleration_matches_the_reference_loop(gpu):
    """ipf_set_newton_thomas against the reference's accelerated loop
    (main.cpp:160-262 with modified_thomas_accel_Q, :194-224) restated on the
    host with the library's primitives -- Assemble = ipf_residual_tangent,
    AssembleResidual = ipf_residual (the tangent stays), Solve =
    ipf_cg_solve -- and the reference's stop test (:239, tolerance 1e-7):
    load steps 1-4 of config 1, the same Newton iteration counts and the same
    u to 1e-10 relative (the host loop's vector updates and dots round in
    torch; measured: the same bits); both end within 1e-7 relative of the
    plain Newton solution (both stop at |r| < 1e-7, |du| < 1e-7; measured
    1.9e-8).  The
    line search and the acceleration together are rejected."""
    import torch
    tol, cg_rtol = 1e-7, 1e-13
    sols = {}
    for mode in ("library", "host", "plain"):
        m, sm, ctx, B, ix, mask = _setup("40x40-q2")
        free = torch.as_tensor(~mask, device=gpu)
        f_unit = ix["rhs_linear"] / wb.normal_stress(wb.NSTEPS)
        neq = sm.sizes()["neq"]
        u = torch.zeros(neq, dtype=torch.float64, device=gpu)
        ctx.set_newton_reference_stop(tol)
        ctx.set_newton_thomas(mode == "library")
        its = []
        for k in range(1, 5):
            ctx.set_rhs_linear(f_unit * wb.normal_stress(k))
            if mode != "host":
                st = ctx.newton(u, rtol=1e-30, maxit=40, cg_rtol=cg_rtol, cg_maxit=40000)
                assert st["converged"] == 1.0, (mode, k, st)
                its.append(int(st["iterations"]))
                continue
            r = torch.empty_like(u)
            d = torch.empty_like(u)
            t = torch.empty_like(u)
            ctx.residual_tangent(u, r)
            alpha, n = 1.0, 0
            for i in range(40):
                ctx.cg_solve(r, d, rtol=cg_rtol, maxit=40000)
                u += alpha * d
                ctx.residual(u, r)
                ctx.cg_solve(r, t, rtol=cg_rtol, maxit=40000)
                du_norm = float(torch.linalg.norm(alpha * d + t))
                num = float(torch.dot(alpha * d, t))
                dem = float(torch.dot(alpha * d, alpha * d))
                alpha += num / dem
                u += t
                ctx.residual_tangent(u, r)
                n = i + 1
                if float(torch.linalg.norm(r[free])) < tol and du_norm < tol:
                    break
            else:
                raise AssertionError(f"host loop did not converge at load step {k}")
            its.append(n)
            ctx.set_update_mem(True)
            ctx.residual(u, r)
            ctx.set_update_mem(False)
        sols[mode] = (u.cpu().numpy(), its)
        print(f"{mode}: Newton iterations per load step {its}", flush=True)
        if mode == "library":
            ctx.set_newton_line_search(4)
            with pytest.raises(ipf.IPFError):
                ctx.newton(u, rtol=1e-30, maxit=2, cg_rtol=cg_rtol, cg_maxit=40000)
    ul, il = sols["library"]
    uh, ih = sols["host"]
    up, _ = sols["plain"]
    print(f"max |u| {np.abs(up).max():.3e}; library - host {np.abs(ul - uh).max():.3e}, "
          f"library - plain {np.abs(ul - up).max():.3e}", flush=True)
    assert il == ih
    assert np.abs(ul - uh).max() <= 1e-10 * np.abs(uh).max()
    assert np.abs(ul - up).max() <= 1e-7 * np.abs(up).max()
