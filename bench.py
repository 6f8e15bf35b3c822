"""Residual throughput of the MI355X integration-point path (one GPU).

A step is one full residual evaluation r = f_lin - sum_e B_e^T (w sigma(B_e u_e))
(TPZElastoPlasticIntPointsStructMatrix::Assemble(rhs)) on the 800x800 Q2
quarter-annulus wellbore (640,000 elements, 5.76 M integration points,
5,126,402 DoF), plain FP64, with u and r resident in HBM, at the converged
Newton state of the 10 load steps (||r||/||r0|| <= 1e-8; the solve -- Newton
with block-Jacobi PCG inner solves, FP64 throughout -- runs first, untimed).  The
data is set up on the GPU from the element corners, and the library's B
format is the geometric one: B and the weights recomputed inside the kernel
from 8 corner coordinates per element, bit-identical to the stored dense and
compact FP64 blocks built from the same geometry.  The plain residual is the
drop-in's: colored assembly, the reference's return map at every point.

Before the W warm-up steps, --settle-ms of untimed back-to-back residuals
bring the GPU to its sustained clocks (config.settle); the timed region is
exactly K residuals.

Prints ONE JSON line on stdout (progress goes to stderr):
  value       Mint-pts/s = points * steps / timed seconds / 1e6
  roofline    the fused per-color residual kernel against HBM: its compulsory
              bytes -- what the B format stores (geometric: 8 corner
              coordinates per element; compact: dN/dx, dN/dy per node and
              point, + a weight per point), a 4-byte DoF index per element
              column, 4 plastic-strain components per point, u read + r
              written once -- over the kernel's average duration from HIP
              events around each launch (dispatch gap removed); `traffic` =
              HBM bytes per launch measured by rocprofv3 PMC (profiles/),
              `traffic_over_algorithmic` its ratio to the compulsory bytes,
              `second_bound` the VALU issue fraction (same profile)
  cpu_baseline the CPU oracle (same algorithm, OpenMP) on the host cores
  residual_paths  the same residual with the patch assembly and / or the
              certified-elastic stress path (ipf_set_residual_assembly /
              ipf_set_residual_mode): time and deviation
  b_format_variants  the same residual with the other B formats (dense and
              compact FP64 and FP32; ipf_set_b_format): time, rate on each
              format's stored bytes, deviation from the FP64 residual
  learned_mixed_batched  the 64 uneven instances at their converged
              mixed-precision Newton state: the plain batched residual and the
              learned + mixed one (predictor-ordered elements, certified-elastic
              path, the faster assembly), speed-up, deviation from the plain
              residual and from the CPU oracle, the predictor's figures and the
              ablation of each ingredient
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "residual Mint-pts/s, plain FP64 vs learned+mixed, at Newton ‖r‖/‖r0‖≤1e-8"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(blocks, npts, neq, fmt="dense-fp64"):
    """Bytes one residual must move: the B data once (dense: every block
    entry; compact: dN/dx, dN/dy per node and point; geometric: 8 corner
    coordinates per element), the weight per point (not for the geometric
    form, which computes it), a 4-byte DoF index per block column, 4
    plastic-strain components per point, and per DoF u and f_lin read once
    and r written once (r = f_lin - sum B^T w sigma)."""
    import numpy as np
    R = blocks["row_sizes"].astype(np.int64)
    C = blocks["col_sizes"].astype(np.int64)
    if fmt == "geometry-fp64":
        return int((64 + 4 * C).sum()) + npts * 32 + neq * (8 + 8 + 8)
    word = 4 if fmt.endswith("fp32") else 8
    nval = (R // 3) * C if fmt.startswith("compact") else R * C
    return int((word * nval + 4 * C).sum()) + npts * (8 + 32) + neq * (8 + 8 + 8)


_ROCTX = None


def roctx_range(name=None):
    """Pushes (name given) or pops a roctx range, which rocprofv3
    --marker-trace records: the timed region can then be selected from a
    kernel trace of the whole run.  A no-op where the roctx library is absent."""
    global _ROCTX
    if _ROCTX is None:
        import ctypes
        _ROCTX = False
        for lib in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
            try:
                _ROCTX = ctypes.CDLL(lib)
                break
            except OSError:
                pass
    if not _ROCTX:
        return
    if name is None:
        _ROCTX.roctxRangePop()
    else:
        _ROCTX.roctxRangePushA(name.encode())


def settle(ctx, u, r, ms):
    """Back-to-back residuals for about `ms` of wall time, untimed: the GPU's
    clocks ramp up under sustained load (the first tens of milliseconds of
    residuals after idle run up to 10 % slower), and a solver's residuals run
    at the sustained clocks.  Returns the number of residuals run."""
    import torch
    n = 0
    t0 = time.perf_counter()
    while (time.perf_counter() - t0) * 1e3 < ms:
        for _ in range(10):
            ctx.residual(u, r)
        torch.cuda.synchronize()
        n += 10
    return n


def time_residuals(ctx, u, r, stream, steps, warmup, roi=None):
    import torch
    for _ in range(warmup):
        ctx.residual(u, r)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    if roi:
        roctx_range(roi)
    start.record(stream)
    for _ in range(steps):
        ctx.residual(u, r)
    end.record(stream)
    torch.cuda.synchronize()
    if roi:
        roctx_range()
    return start.elapsed_time(end)


def kernel_name(fmt, shape, pairs=True, fast=False):
    """Name rocprofv3 shows for the non-committing residual kernel of a B
    format and element shape (pairs: one gathered index per node, taken when
    every element numbers its nodes' DoFs as (x, y) pairs, as the wellbore
    meshes do; fast: the certified-elastic stress mode)."""
    import intpointsfem_amd as ipf
    tf = lambda b: "true" if b else "false"
    return f"void k_residual_fmt<{shape}, {ipf.B_FORMATS.index(fmt)}, false, {tf(pairs)}, {tf(fast)}>"


def launch_roofline(ctx, u, r, nrep, blocks, npts, neq, fmt, step_ms=None):
    """Per-launch HBM rates of the residual kernels: each launch is credited
    with its points' share of the compulsory bytes of what the format reads
    (algorithmic_bytes).  Events around every launch add the dispatch gap
    that back-to-back launches hide; with `step_ms` (the back-to-back residual
    time) that gap -- (sum of the intervals that hold a kernel - step_ms) /
    their number -- is taken off each kernel interval, so the durations are
    those of the timed region.  The interval before the kernels (the f_lin
    copy) holds no kernel when the launches read f_lin themselves (colored
    assembly with every DoF in some element): it is left out of the gap and
    reported as measured (the library reports an empty interval as 0)."""
    avg, pts = launch_times(ctx, u, r, nrep)
    kern = pts > 0  # the intervals of the residual kernels proper
    first_real = float(avg[0]) > 0.0  # a copy / refresh kernel ran in interval 0 (else exactly 0)
    has = kern.copy()
    has[0] = first_real
    tail = np.zeros(len(avg), dtype=bool)
    tail[1:] = ~kern[1:]  # e.g. the patch fix-up: a kernel without points of its own
    has |= tail
    gap = 0.0
    if step_ms is not None:
        gap = max(0.0, (float(avg[has].sum()) - step_ms) / int(has.sum()))
    isolated = float(avg[kern].mean())
    adj = np.where(has, avg - gap, avg)
    main_ms = adj[kern]
    share = pts[kern] / npts
    sto = algorithmic_bytes(blocks, npts, neq, fmt) * share
    t = main_ms.sum() / 1e3
    return {"copy_ms": float(adj[0]), "color_ms": main_ms, "other_ms": [float(x) for x in adj[tail]],
            "avg_launch_ms": float(main_ms.mean()), "avg_launch_ms_isolated": isolated, "dispatch_gap_ms": gap,
            "stored_bytes_per_launch": float(sto.mean()), "stored_bytes": algorithmic_bytes(blocks, npts, neq, fmt),
            "achieved_stored": float(sto.sum() / t / 1e9)}


def launch_times(ctx, u, r, nrep):
    """Average per-launch durations (HIP events on the launch stream): index 0
    is the f_lin copy (empty when the color launches read f_lin themselves),
    then one entry per color kernel; and the points each launch covers."""
    acc = None
    pts = None
    for _ in range(nrep):
        ms, p = ctx.residual_timed(u, r)
        acc = ms if acc is None else acc + ms
        pts = p
    return acc / nrep, pts


def event_ms(fn, stream, reps=10):
    import torch
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps




def profile_figures(workload, kname):
    """HBM bytes per launch (rocprofv3 PMC: the L2's DRAM read and write
    requests, checked on a copy kernel of known size) and VALU issue fraction
    of kernel `kname` for `workload`, from the committed profile summary
    (profiles/); (None, None) when the profile does not cover this kernel."""
    tfile = os.path.join(ROOT, "profiles", f"k_residual_{workload}_traffic.json")
    if not os.path.exists(tfile):
        return None, None
    with open(tfile) as f:
        tk = json.load(f)["kernels"]
    if kname not in tk:
        return None, None
    k = tk[kname]
    return (int(k["total_bytes"]) if "total_bytes" in k else None), k.get("valu_busy")


def roofline_entry(roof, kname, traffic, valu_busy, npts):
    """The line's roofline object: compulsory bytes per launch over the
    launch's average duration, against HBM; the measured HBM traffic per
    launch (rocprofv3 PMC) and its ratio to the compulsory bytes; the VALU
    issue fraction as the second bound."""
    alg = roof["stored_bytes_per_launch"]
    return {"bound": "hbm", "achieved": round(roof["achieved_stored"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(roof["achieved_stored"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
            "algorithmic_bytes_per_launch": int(alg),
            "algorithmic_bytes_per_point": round(roof["stored_bytes"] / npts, 1),
            "avg_launch_ms": round(roof["avg_launch_ms"], 5),
            "avg_launch_ms_isolated": round(roof["avg_launch_ms_isolated"], 5),
            "dispatch_gap_ms": round(roof["dispatch_gap_ms"], 5),
            "traffic_over_algorithmic": None if traffic is None else round(traffic / alg, 3),
            "measured_traffic_rate": None if traffic is None else round(traffic / (roof["avg_launch_ms"] / 1e3) / 1e9, 1),
            # 4 cycles x VALU instructions / (1024 SIMDs x 2.4 GHz x duration), rocprofv3 PMC (profiles/)
            "valu_busy": None if valu_busy is None else round(valu_busy, 3),
            "second_bound": None if valu_busy is None else {"bound": "valu", "frac": round(valu_busy, 3)}}


def residual_paths(ctx, u, r, stream, steps, warmup, npts, r_ref, fint, assemblies=("colored", "patch")):
    """The non-committing residual in every (assembly, stress mode) the
    library offers, at the same state and u: time per residual, rate, and
    deviation from the reference-order exact residual r_ref."""
    out = {}
    for asm in assemblies:
        try:
            ctx.set_residual_assembly(asm)
        except Exception as e:  # the patch assembly needs the geometric setup
            out[asm] = {"unavailable": str(e)}
            continue
        for mode in ("exact", "certified-elastic"):
            ctx.set_residual_mode(mode)
            tms = time_residuals(ctx, u, r, stream, steps, warmup)
            ctx.residual(u, r)
            ctx.synchronize()
            dev = float(np.abs(r.cpu().numpy() - r_ref).max()) / fint
            out[f"{asm}/{mode}"] = {"ms_per_step": round(tms / steps, 5),
                                    "value": round(npts * steps / (tms / 1e3) / 1e6, 2),
                                    "max_dev_vs_exact_colored": dev}
            log(f"[bench] residual {asm}/{mode}: {tms / steps:.4f} ms, {out[f'{asm}/{mode}']['value']:.1f} "
                f"Mint-pts/s, deviation {dev:.2e} of max|f_int|")
    ctx.set_residual_assembly("colored")
    ctx.set_residual_mode("exact")
    return out


def train_predictor(ctx, u, npts, dev, seed=0, steps=3000):
    """The branch predictor trained on the exact return map's pairs at this
    state (a seeded 90 / 10 split), on the GPU.  Returns (mlp, features,
    stats) with the held-out accuracy and the inference time over all points."""
    import torch

    from intpointsfem_amd import learned as L
    t0 = time.time()
    feat, br, tg = L.point_features(ctx, u, npts, dev)
    g = torch.Generator(device="cpu").manual_seed(seed)
    held = (torch.rand(npts, generator=g) < 0.1).to(dev)
    tr = torch.nonzero(~held).squeeze(1)
    te = torch.nonzero(held).squeeze(1)
    mlp = L.new_mlp(seed=seed)
    L.train(mlp, feat.index_select(0, tr), br.index_select(0, tr), tg.index_select(0, tr), steps=steps,
            batch=4096, lr=3e-3, seed=seed)
    torch.cuda.synchronize()
    train_s = time.time() - t0
    pred = torch.zeros(npts, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    pms = event_ms(lambda: mlp.predict(feat, branch=pred, n=npts), stream)
    acc = float((pred.index_select(0, te) == br.index_select(0, te)).float().mean())
    major = float(torch.bincount(br.index_select(0, te).long(), minlength=5).max()) / max(1, te.shape[0])
    stats = {"pairs": npts, "train_seconds": round(train_s, 1), "held_out_accuracy": round(acc, 4),
             "majority_class": round(major, 4), "predict_ms": round(pms, 4),
             "predict_ms_per_5.76M_points": round(pms * 5.76e6 / npts, 4),
             "predict_rate_Gpts_s": round(npts / pms / 1e6, 2)}
    log(f"[bench] predictor: {train_s:.1f} s to train on {npts} pairs, held-out accuracy {acc:.4f} (majority "
        f"{major:.4f}); inference {pms:.4f} ms for {npts} points")
    return mlp, pred, stats


def element_keys(pred, row_sizes):
    """Per element: 1 when the predictor puts any of its points off the
    elastic branch (the element then cannot join a certified-elastic tile)."""
    nip = (np.asarray(row_sizes) // 3).astype(np.int64)
    p = pred.cpu().numpy()
    first = np.concatenate([[0], np.cumsum(nip)[:-1]])
    return (np.maximum.reduceat((p != 0).astype(np.int32), first) > 0).astype(np.int32)


def learned_mixed_batched(args, dev):
    """The learned + mixed-precision batched residual on the 64 uneven
    instances (tests/golden/batch64_instances.json, drawn with seed 0):

      * state (--batch-state): lame (default) -- each instance's plastic
        state committed at twice its linear-elastic (Lame) displacement, the
        residuals at three times it; newton -- 10 load steps to each
        instance's final wellbore pressure by the segmented Newton with FP32
        tangent products in the CG and FP64 residuals (iterative
        refinement), ||r_k|| / ||r0_k|| <= newton_rtol per instance, committed;
      * plain: the FP64 residual of the drop-in (colored assembly, exact
        return map), natural element order;
      * learned + mixed: the predictor (trained on this state's pairs) orders
        each color's / patch's elements so that predicted-elastic elements
        share tiles, and the certified-elastic path skips the spectral
        decomposition of every point it proves elastic; the tile shape
        (colored tiles or patches of 2 / 4 / 8 wavefront tiles) is the one
        ipf_autotune_assembly picks at this state;
      * deviation from the plain residual, and from the CPU oracle on whole
        instances within --cpu-seconds.
    Returns (dict for the JSON line, batch, u, plain_ms)."""
    import torch

    from intpointsfem_amd import batch
    from intpointsfem_amd import wellbore as wb
    table = batch.draw_instances(64, seed=0)
    t0 = time.time()
    meshes = [batch.instance_mesh(i) for i in table]
    B = batch.Batch(meshes, device=0)
    ctx = B.ctx
    npts, neq = B.npts, B.neq
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    B.set_dirichlet_mask(B.dirichlet_mask())
    log(f"[bench] batch64: {len(B.row_sizes)} elements, {npts} points, neq {neq}; setup {time.time() - t0:.1f}s")
    newton = None
    if args.batch_state == "newton":
        # each instance on its own: 10 load steps to its final pressure (its
        # load is linear in the step), Newton with the mixed tangent (FP32
        # products in the CG, FP64 residuals) to its own criterion; then the
        # converged u and committed state go into the batch (the batch's
        # equations and points are the instances' in table order).  One CG
        # over the whole batch would take its relative tolerance over all
        # instances at once: the instances with small residuals then get
        # updates that are not accurate for them, and the batched Newton
        # stalls from load step 2 on this table.
        import intpointsfem_amd as ipf
        tfmt = args.batch_tangent_format
        t1 = time.time()
        ue, eps, als, its = [], [], [], []
        cgs, worst, conv = 0, 0.0, True
        for i, m in enumerate(meshes):
            smi = ipf.StructMatrix(m, device=0)
            ci = smi.context()
            ci.set_stream(stream.cuda_stream)
            ci.set_dirichlet_mask(m.dirichlet_mask())
            ci.set_b_format(ci.b_format()[0], tangent=tfmt)
            ci.set_newton_line_search(args.batch_line_search)
            fi = smi.export_indexes()["rhs_linear"]
            ui = torch.zeros(smi.sizes()["neq"], dtype=torch.float64, device=dev)
            ni = 0
            for k in range(1, wb.NSTEPS + 1):
                ci.set_rhs_linear(fi * (k / wb.NSTEPS))
                st = ci.newton(ui, rtol=args.newton_rtol, maxit=80, cg_rtol=args.cg_rtol, cg_maxit=200000)
                ni += int(st["iterations"])
                cgs += int(st["cg_iterations"])
                worst = max(worst, st["rel_residual"])
                conv = conv and st["converged"] == 1.0
                if st["converged"] != 1.0 or k == wb.NSTEPS:
                    log(f"[bench]   batch64 instance {i} (N={table[i]['N']}, p={table[i]['p']}) load step {k}: "
                        f"newton {int(st['iterations'])}, ||r||/||r0|| {st['rel_residual']:.2e}, "
                        f"{time.time() - t1:.1f}s")
            ep_i, al_i, _, _ = ci.get_state()
            ue.append(ui.cpu().numpy())
            eps.append(ep_i)
            als.append(al_i)
            its.append(ni)
            del ci, smi
        u = torch.as_tensor(np.concatenate(ue), dtype=torch.float64, device=dev)
        ctx.set_state(np.concatenate(eps), np.concatenate(als))
        ctx.synchronize()
        newton = {"tangent_format": tfmt, "line_search_halvings": args.batch_line_search,
                  "newton_iterations_per_instance": its, "converged": conv,
                  "worst_final_rel_residual": worst, "cg_iterations": cgs,
                  "solve_seconds": round(time.time() - t1, 1),
                  "load": f"{wb.NSTEPS} load steps to each instance's final pressure, each instance on its own"}
        log(f"[bench] batch64 newton: {sum(its)} Newton iterations, {cgs} CG, converged {conv}, "
            f"{newton['solve_seconds']} s")
        state = (f"Newton solution of load step {wb.NSTEPS} of every instance ({tfmt} tangent products in the CG, "
                 f"FP64 residuals), committed: " + (f"converged, ||r||/||r0|| <= {args.newton_rtol:g} per instance"
                                                    if conv else "NOT converged (see newton)"))
    else:
        # each instance's linear-elastic (Lame) displacement: 2x committed, residuals at 3x
        us = [wb.lame_displacement(m.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, i["pw_final"], wb.SIGMA0),
                                   E_=i["E"], nu=i["nu"]) for i, m in zip(table, meshes)]
        u1 = torch.as_tensor(np.concatenate([2 * x for x in us]), device=dev)
        u = torch.as_tensor(np.concatenate([3 * x for x in us]), device=dev)
        r1 = torch.zeros(neq, dtype=torch.float64, device=dev)
        ctx.set_update_mem(True)
        ctx.residual(u1, r1)
        ctx.set_update_mem(False)
        ctx.synchronize()
        del u1, r1
        state = ("plastic state committed at 2x each instance's linear-elastic (Lame) displacement, residuals at "
                 "3x (--batch-state newton: the converged 10-step Newton state)")
    _, _, mt, br = ctx.get_state()
    r = torch.zeros(neq, dtype=torch.float64, device=dev)
    settle(ctx, u, r, args.settle_ms)
    plain_total = time_residuals(ctx, u, r, stream, args.steps, args.warmup, roi="timed plain residuals")
    plain_ms = plain_total / args.steps
    ctx.residual(u, r)
    ctx.synchronize()
    r_plain = r.cpu().numpy()
    fint = float(np.abs(r_plain - B.rhs_linear).max())
    ablation = {}

    def timed(label):
        tms = time_residuals(ctx, u, r, stream, args.steps, args.warmup) / args.steps
        ctx.residual(u, r)
        ctx.synchronize()
        dev_ = float(np.abs(r.cpu().numpy() - r_plain).max()) / fint
        ablation[label] = {"ms_per_step": round(tms, 5), "speedup_vs_plain": round(plain_ms / tms, 3),
                           "max_dev_vs_plain": dev_}
        log(f"[bench] batch64 {label}: {tms:.4f} ms ({plain_ms / tms:.3f}x plain), deviation {dev_:.2e}")
        return tms

    for asm in ("colored", "patch"):
        ctx.set_residual_assembly(asm)
        ctx.set_residual_mode("exact")
        if asm == "patch":
            timed("patch/exact, index order")
        ctx.set_residual_mode("certified-elastic")
        timed(f"{asm}/certified-elastic, index order")
    ctx.set_residual_assembly("colored")
    ctx.set_residual_mode("exact")
    # the predictor: trained on this state's pairs, then the element order
    mlp, pred, pstats = train_predictor(ctx, u, npts, dev)
    keys = element_keys(pred, B.row_sizes)
    t2 = time.time()
    B.relayout(keys)
    ctx.set_stream(stream.cuda_stream)
    relayout_s = time.time() - t2
    # the tile shape: colored tiles or patches of 2 / 4 / 8 wavefront tiles, by the autotuner
    ctx.set_residual_mode("certified-elastic")
    tune = ctx.autotune_assembly(u, reps=max(5, min(args.steps, 50)))
    log(f"[bench] batch64 assembly autotune (certified-elastic, predicted order): "
        f"{', '.join(f'{k} {v:.4f} ms' for k, v in tune['times_ms'].items())} -> {tune['selected']}")
    best = (tune["selected"], timed(f"{tune['selected']}/certified-elastic, predicted order"))
    settle(ctx, u, r, args.settle_ms)
    lm_total = time_residuals(ctx, u, r, stream, args.steps, args.warmup, roi="timed learned+mixed residuals")
    lm_ms = lm_total / args.steps
    ctx.residual(u, r)
    ctx.synchronize()
    r_lm = r.cpu().numpy()
    dev_plain = float(np.abs(r_lm - r_plain).max()) / fint
    # the CPU oracle on whole instances (in table order) within the time bound
    oracle_dev, checked = None, 0
    if not args.no_cpu_baseline:
        import intpointsfem_amd as ipf
        from oracle import oracle as O
        ep, al, _, _ = ctx.get_state()
        uh = u.cpu().numpy()
        spent, worst = 0.0, 0.0
        for i, m in enumerate(meshes):
            sm = ipf.StructMatrix(m, device=-1)
            b = sm.export_blocks()
            ix = sm.export_indexes()
            s = sm.sizes()
            Bo = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                          b["storage"])
            d = m.material()
            e0, e1 = B.eq_offsets[i], B.eq_offsets[i + 1]
            p0, p1 = B.pt_offsets[i], B.pt_offsets[i + 1]
            t = time.time()
            ro, *_ = O.residual(Bo, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"],
                                d["lambda"], d["mu"], d["phi"], d["psi"], d["cohesion"], uh[e0:e1],
                                f_lin=ix["rhs_linear"], plastic_strain6=ep[6 * p0:6 * p1], alpha=al[p0:p1],
                                nthreads=max(1, min(16, os.cpu_count() or 1)))
            spent += time.time() - t
            worst = max(worst, float(np.abs(r_lm[e0:e1] - ro).max() / np.abs(ro - ix["rhs_linear"]).max()))
            checked += 1
            if spent >= args.cpu_seconds:
                break
        oracle_dev = worst
        log(f"[bench] batch64 learned+mixed vs the CPU oracle on {checked} instances: {worst:.2e} of max|f_int|")
    ctx.set_residual_assembly("colored")
    ctx.set_residual_mode("exact")
    value = npts * args.steps / (lm_total / 1e3) / 1e6
    log(f"[bench] batch64 learned+mixed ({best[0]} assembly): {lm_ms:.4f} ms, {value:.1f} Mint-pts/s, "
        f"{plain_ms / lm_ms:.3f}x plain ({plain_ms:.4f} ms)")
    out = {"workload": "64 uneven wellbore instances (seed-0 table)", "state": state,
           "value": round(value, 2), "unit": "Mint-pts/s", "ms_per_step": round(lm_ms, 5),
           "speedup_vs_plain": round(plain_ms / lm_ms, 3),
           "plain": {"ms_per_step": round(plain_ms, 5), "value": round(npts * args.steps / (plain_total / 1e3) / 1e6, 2),
                     "path": "colored assembly, exact return map, index order (the drop-in's residual)"},
           "path": {"assembly": best[0], "residual_mode": "certified-elastic", "order": "predicted branch",
                    "assembly_autotune_ms": {k: round(v, 5) for k, v in tune["times_ms"].items()}},
           "max_dev_vs_plain": dev_plain, "max_dev_vs_oracle": oracle_dev, "oracle_instances_checked": checked,
           "int_points": npts, "neq": neq, "newton": newton,
           "branches": {n: int((br == i).sum()) for i, n in enumerate(("elastic", "main", "right_edge", "left_edge",
                                                                       "apex"))},
           "predictor": dict(pstats, elements_keyed_plastic=int(keys.sum()), elements=int(len(keys)),
                             relayout_seconds=round(relayout_s, 1)),
           "ablation": ablation, "steps": args.steps, "warmup": args.warmup}
    return out, B, u, plain_ms


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed back-to-back residuals before the warm-up steps, in ms of wall time")
    ap.add_argument("--mesh", default="800x800-q2",
                    help="800x800-q2 (default), 800x800-q3, 40x40-q2, or batch64 (the 64 uneven instances only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-formats", action="store_true", help="skip the compact / FP32 B format variants")
    ap.add_argument("--no-batch", action="store_true", help="skip the learned + mixed batched measurement")
    ap.add_argument("--no-predictor", action="store_true",
                    help="skip training / timing the branch predictor on this mesh's state")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the library's default B format instead of the autotuner's choice")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound on the CPU baseline's timed work")
    ap.add_argument("--state", choices=("newton", "lame"), default="newton",
                    help="newton (default): the converged 10-step Newton solution (||r||/||r0|| <= --newton-rtol) "
                         "with its committed plastic state, the state the metric is quoted at; lame: the "
                         "thick-cylinder displacement of the last load step with zero plastic strain (no solve)")
    ap.add_argument("--newton-rtol", type=float, default=1e-8)
    ap.add_argument("--cg-rtol", type=float, default=1e-1,
                    help="inner block-Jacobi PCG tolerance of each Newton step (the Newton loop converges linearly "
                         "with the reference's tangent, so a loose inner solve costs few extra Newton steps)")
    ap.add_argument("--save-state", default=None, metavar="FILE",
                    help="--state newton: write the converged state (u, plastic strain, alpha) to FILE (.npz)")
    ap.add_argument("--load-state", default=None, metavar="FILE",
                    help="--state newton: start from a state written by --save-state for the same mesh instead of "
                         "solving (profiler passes over the same state)")
    ap.add_argument("--solve-tangent-format", default=None,
                    help="B format of the CG matvecs of the --state newton solve (default: the library's, "
                         "geometric FP64; the residual stays FP64 whatever this is)")
    ap.add_argument("--batch-state", choices=("lame", "newton"), default="lame",
                    help="state of the batched measurement: a committed plastic state from each instance's "
                         "linear-elastic displacement (fast), or 10 load steps of Newton with a line search for "
                         "every instance, instance by instance (over 15 minutes; with the inexact inner solves of "
                         "--cg-rtol some low-cohesion instances do not converge, which the line reports)")
    ap.add_argument("--batch-line-search", type=int, default=8,
                    help="step halvings of the Newton line search in --batch-state newton (0: full steps)")
    ap.add_argument("--batch-tangent-format", default="compact-fp32",
                    help="B format of the CG matvecs of the batch's Newton solve (FP32: mixed precision)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the path runs on a single "
              "MI355X; multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    import torch

    import intpointsfem_amd as ipf
    from intpointsfem_amd import wellbore as wb

    dev = torch.device("cuda:0")
    if args.mesh == "batch64":
        lm, B, u, plain_ms = learned_mixed_batched(args, dev)
        out = {"metric": METRIC, "value": lm["plain"]["value"], "unit": "Mint-pts/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": lm["plain"]["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic",
               "config": {"workload": "plain FP64 batched residual, 64 uneven wellbore instances (seed-0 table)",
                          "state": lm["state"], "int_points": B.npts,
                          "neq": B.neq, "instances": B.n},
               "roofline": None, "cpu_baseline": None, "learned_mixed_batched": lm}
        print(json.dumps(out), flush=True)
        return

    t0 = time.time()
    mesh = wb.make_mesh(args.mesh)
    info = mesh.info()
    log(f"[bench] mesh {args.mesh}: {info['nelem']} elements, neq {info['neq']}")
    sm = ipf.StructMatrix(mesh, device=0)
    ctx = sm.context()
    sizes = sm.sizes()
    npts, neq = sizes["npts"], sizes["neq"]
    log(f"[bench] setup + upload {time.time() - t0:.1f}s; {npts} points, {sizes['ncolor']} colors, "
        f"{ctx.info()['device_bytes'] / 1e9:.2f} GB on device")
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)

    r = torch.zeros(neq, dtype=torch.float64, device=dev)
    newton_info = None
    if args.state == "lame":
        u_host = wb.lame_displacement(mesh.nodes(), wb.normal_stress(wb.NSTEPS), scale=1.0)
        u = torch.as_tensor(u_host, dtype=torch.float64, device=dev)
        ctx.set_update_mem(True)
        ctx.residual(u, r)
        ctx.set_update_mem(False)
        ctx.synchronize()
        _, _, mt, br = ctx.get_state()
        ctx.set_state(np.zeros(6 * npts))
        state_desc = "u = linear-elastic thick-cylinder displacement of load step 10, zero plastic strain"
    elif args.load_state:
        z = np.load(args.load_state, allow_pickle=False)
        if str(z["mesh"]) != args.mesh or int(z["neq"]) != neq:
            raise SystemExit(f"{args.load_state} holds a state of {z['mesh']} (neq {int(z['neq'])}), not {args.mesh}")
        u_host = z["u"]
        u = torch.as_tensor(u_host, dtype=torch.float64, device=dev)
        ctx.set_state(z["ep"], z["alpha"])
        ctx.synchronize()
        newton_info = json.loads(str(z["newton_info"]))
        newton_info["loaded_from"] = os.path.basename(args.load_state)
        br = z["branch"]
        state_desc = (f"converged Newton solution of load step 10 (||r||/||r0|| <= {args.newton_rtol:g}), "
                      "committed plastic strain of the 10 load steps (loaded from --save-state output)")
    else:
        # 10 load steps of Newton + block-Jacobi PCG on the GPU; the state is committed after each
        ctx.set_dirichlet_mask(mesh.dirichlet_mask())
        fmt_default = ctx.b_format()
        solve_tf = args.solve_tangent_format or fmt_default[1]
        ctx.set_b_format(fmt_default[0], tangent=solve_tf)
        f_unit = sm.export_indexes()["rhs_linear"] / wb.normal_stress(wb.NSTEPS)
        u = torch.zeros(neq, dtype=torch.float64, device=dev)
        t1 = time.time()
        steps = []
        ctx.set_newton_monitor(lambda it, rel, rn, cg: log(f"[bench]   newton {it}: ||r||/||r0|| {rel:.2e}, cg {cg}, "
                                                              f"{time.time() - t1:.1f}s") if it % 5 == 0 else None)
        for k in range(1, wb.NSTEPS + 1):
            ctx.set_rhs_linear(f_unit * wb.normal_stress(k))
            st = ctx.newton(u, rtol=args.newton_rtol, maxit=60, cg_rtol=args.cg_rtol, cg_maxit=200000)
            steps.append(st)
            log(f"[bench] load step {k}: newton {int(st['iterations'])}, |r|/|r0| {st['rel_residual']:.2e}, "
                f"cg {int(st['cg_iterations'])}, {time.time() - t1:.1f}s")
        ctx.set_newton_monitor(None)
        ctx.set_b_format(*fmt_default)
        u_host = u.cpu().numpy()
        _, _, mt, br = ctx.get_state()
        newton_info = {"solve_tangent_format": solve_tf, "load_steps": wb.NSTEPS,
                       "newton_iterations": [int(s["iterations"]) for s in steps],
                       "final_rel_residual": steps[-1]["rel_residual"],
                       "converged": all(s["converged"] == 1.0 for s in steps),
                       "cg_iterations": int(sum(s["cg_iterations"] for s in steps)),
                       "solve_seconds": round(time.time() - t1, 1)}
        state_desc = (f"converged Newton solution of load step 10 (||r||/||r0|| <= {args.newton_rtol:g}), "
                      "committed plastic strain of the 10 load steps")
        if args.save_state:
            ep_s, al_s, _, _ = ctx.get_state()
            np.savez(args.save_state, mesh=args.mesh, neq=neq, u=u_host, ep=ep_s, alpha=al_s, branch=br,
                     newton_info=json.dumps(newton_info))
            log(f"[bench] state saved to {args.save_state}")
    branch_counts = {ipf.BRANCH_NAMES[i]: int((br == i).sum()) for i in range(5)}
    log(f"[bench] branches: {branch_counts}")
    ep_state, al_state, _, _ = ctx.get_state()  # the committed state the timed residuals start from

    tune = None
    if not args.no_autotune:
        # the fastest exact FP64 B format for these data (all give the same bits)
        tune = ctx.autotune(u, reps=10)
        log(f"[bench] autotune: {', '.join(f'{k} {v:.4f} ms' for k, v in tune['residual'].items())} -> "
            f"{tune['selected'][0]}")
        tune = {"residual_ms": {k: round(v, 5) for k, v in tune["residual"].items()}, "selected": tune["selected"][0]}
    fmt0 = ctx.b_format()[0]  # the autotuner's choice (geometric FP64 after the device setup without it)
    shape = {1: 0, 2: 1, 3: 2}[info["porder"]]
    nsettle = settle(ctx, u, r, args.settle_ms)
    total_ms = time_residuals(ctx, u, r, stream, args.steps, args.warmup, roi="timed residuals")
    ms_per_step = total_ms / args.steps
    value = npts * args.steps / (total_ms / 1e3) / 1e6
    log(f"[bench] {args.steps} residuals ({fmt0} B): {ms_per_step:.4f} ms/residual, {value:.1f} Mint-pts/s")

    # per-launch durations (HIP events on the launch stream) for the roofline
    nrep = max(3, min(args.steps, 20))
    blocks = sm.export_blocks()
    roof = launch_roofline(ctx, u, r, nrep, blocks, npts, neq, fmt0, step_ms=ms_per_step)
    # one untimed ipf_residual_tangent: its k_copy (f_lin into r, neq doubles
    # read and written) is the known-size dispatch that checks the PMC byte
    # counters of a profiled run (the residual itself needs no copy)
    ctx.residual_tangent(u, torch.empty_like(r))
    ctx.synchronize()
    kname = kernel_name(fmt0, shape)
    traffic, valu_busy = profile_figures(args.mesh, kname)
    roofline = roofline_entry(roof, f"{kname} (one launch per element color)", traffic, valu_busy, npts)
    log(f"[bench] copy interval {roof['copy_ms']:.4f} ms; color kernels "
        f"{', '.join(f'{x:.4f}' for x in roof['color_ms'])} ms (dispatch gap {roof['dispatch_gap_ms']:.4f} ms "
        f"removed); {roof['stored_bytes'] / npts:.1f} compulsory B/pt -> {roof['achieved_stored']:.0f} GB/s")

    ctx.residual(u, r)
    ctx.synchronize()
    r64 = r.cpu().numpy()
    fint = float(np.abs(r64 - sm.export_indexes()["rhs_linear"]).max())
    paths = residual_paths(ctx, u, r, stream, args.steps, args.warmup, npts, r64, fint)
    predictor = None
    if not args.no_predictor:
        # the branch predictor on this state: its inference kernel over all the points
        _, _, predictor = train_predictor(ctx, u, npts, dev)

    # the other B formats (same state, same u): timing, roofline on their own
    # stored bytes, and the deviation from the FP64 residual
    variants = {}
    if not args.no_formats:
        for fmt in ipf.B_FORMATS:
            if fmt == fmt0:
                continue
            ctx.set_b_format(fmt)
            tms = time_residuals(ctx, u, r, stream, args.steps, args.warmup)
            ctx.residual(u, r)
            ctx.synchronize()
            err = float(np.abs(r.cpu().numpy() - r64).max()) / fint
            fr = launch_roofline(ctx, u, r, nrep, blocks, npts, neq, fmt, step_ms=tms / args.steps)
            variants[fmt] = {"ms_per_step": round(tms / args.steps, 5),
                             "value": round(npts * args.steps / (tms / 1e3) / 1e6, 2),
                             "kernel": kernel_name(fmt, shape),
                             "avg_launch_ms": round(fr["avg_launch_ms"], 5),
                             "stored_bytes_per_pt": round(fr["stored_bytes"] / npts, 1),
                             "achieved_on_stored_bytes": round(fr["achieved_stored"], 1),
                             "frac_on_stored_bytes": round(fr["achieved_stored"] / HBM_PEAK_GBS, 4),
                             "max_dev_vs_fp64": err}
            log(f"[bench] {fmt}: {tms / args.steps:.4f} ms/residual, {variants[fmt]['value']:.1f} Mint-pts/s, "
                f"{fr['achieved_stored']:.0f} GB/s on its stored bytes, "
                f"|r - r_fp64| / max|f_int| = {err:.2e}")
        ctx.set_b_format(fmt0)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        ix = sm.export_indexes()
        B = O.Blocks(blocks["row_sizes"], blocks["col_sizes"], blocks["matrix_position"], blocks["row_first"],
                     blocks["col_first"], blocks["storage"])
        mat = mesh.material()
        nthreads = max(1, min(16, os.cpu_count() or 1))

        def cpu_rates(threads, seconds, max_reps):
            """Mint-pts/s of each full oracle residual, repeated until `seconds` (at least one)."""
            rates, spent = [], 0.0
            while spent < seconds and len(rates) < max_reps:
                t = time.time()
                O.residual(B, ix["dof_indexes"], ix["color_indexes"], sizes["ncolor"], neq, ix["weight"],
                           mat["lambda"], mat["mu"], mat["phi"], mat["psi"], mat["cohesion"], u_host,
                           f_lin=ix["rhs_linear"], plastic_strain6=ep_state, alpha=al_state, nthreads=threads)
                dt = time.time() - t
                spent += dt
                rates.append(npts / dt / 1e6)
            return rates, spent

        rates, el = cpu_rates(nthreads, args.cpu_seconds, 50)
        rates1, el1 = cpu_rates(1, min(args.cpu_seconds, 8.0), 5)
        cpu_value = float(np.median(rates))
        cpu = {"value": round(cpu_value, 3), "unit": "Mint-pts/s", "cores": nthreads, "kind": "port",
               "sample": f"median of {len(rates)} full residual(s) of the same {args.mesh} state ({npts} points) "
                         f"by the C oracle with {nthreads} OpenMP threads, {el:.1f} s (host: {os.cpu_count()} "
                         f"logical CPUs, 16 of them this job's share)",
               "single_thread": {"value": round(float(np.median(rates1)), 3), "reps": len(rates1),
                                 "seconds": round(el1, 1)}}
        log(f"[bench] cpu oracle: {cpu_value:.2f} Mint-pts/s on {nthreads} threads ({len(rates)} reps), "
            f"{cpu['single_thread']['value']:.2f} on 1 thread ({len(rates1)} reps)")
        del B

    lm = None
    if not args.no_batch:
        del sm, ctx
        torch.cuda.synchronize()
        lm, _, _, _ = learned_mixed_batched(args, dev)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mint-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 residual, {args.mesh} quarter-annulus wellbore",
                   "b_format": fmt0, "autotune": tune,
                   "settle": {"ms": args.settle_ms, "residuals": nsettle},
                   "elements": sizes["nblocks"], "int_points": npts, "neq": neq, "colors": sizes["ncolor"],
                   "state": state_desc, "branches": branch_counts, "newton": newton_info},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "residual_paths": paths,
        "predictor": predictor,
        "b_format_variants": variants,
        "learned_mixed_batched": lm,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
