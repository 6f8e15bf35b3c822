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
compact FP64 blocks built from the same geometry.

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
              read-modify-written once -- over the kernel's average duration
              from HIP events around each launch; `traffic` = HBM bytes per
              launch measured by rocprofv3 (profiles/), `dense_fp64_bytes_rate`
              = the rate on the dense blocks' byte count, bytes this kernel
              does not move
  cpu_baseline the CPU oracle (same algorithm, OpenMP) on the host cores
  b_format_variants  the same residual with the other B formats (dense and
              compact FP64 and FP32; ipf_set_b_format): time, rate on each
              format's stored bytes, deviation from the FP64 residual
"""
import argparse
import json
import os
import sys
import time

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
    plastic-strain components per point, u read and r read-modify-written
    once."""
    import numpy as np
    R = blocks["row_sizes"].astype(np.int64)
    C = blocks["col_sizes"].astype(np.int64)
    if fmt == "geometry-fp64":
        return int((64 + 4 * C).sum()) + npts * 32 + neq * (8 + 8)
    word = 4 if fmt.endswith("fp32") else 8
    nval = (R // 3) * C if fmt.startswith("compact") else R * C
    return int((word * nval + 4 * C).sum()) + npts * (8 + 32) + neq * (8 + 8)


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


def kernel_name(fmt, shape, pairs=True):
    """Name rocprofv3 shows for the non-committing residual kernel of a B
    format and element shape (pairs: one gathered index per node, taken when
    every element numbers its nodes' DoFs as (x, y) pairs, as the wellbore
    meshes do)."""
    import intpointsfem_amd as ipf
    return f"void k_residual_fmt<{shape}, {ipf.B_FORMATS.index(fmt)}, false, {'true' if pairs else 'false'}>"


def launch_roofline(ctx, u, r, nrep, blocks, npts, neq, fmt, step_ms=None):
    """Per-launch HBM rates of the color kernels: each launch is credited with
    its points' share of the plain FP64 algorithmic bytes (dense B) and of the
    bytes the format actually stores.  Events around every launch add the
    dispatch gap that back-to-back launches hide; with `step_ms` (the
    back-to-back residual time) that gap, (sum of the per-launch times -
    step_ms) / launches, is taken off each launch, so the durations are those
    of the timed region."""
    avg, pts = launch_times(ctx, u, r, nrep)
    gap = 0.0
    if step_ms is not None:
        gap = max(0.0, (float(avg.sum()) - step_ms) / len(avg))
    isolated = float(avg[1:].mean())
    avg = avg - gap
    color_ms = avg[1:]
    share = pts[1:] / npts
    alg = algorithmic_bytes(blocks, npts, neq) * share
    sto = algorithmic_bytes(blocks, npts, neq, fmt) * share
    t = color_ms.sum() / 1e3
    return {"copy_ms": float(avg[0]), "color_ms": color_ms, "avg_launch_ms": float(color_ms.mean()),
            "avg_launch_ms_isolated": isolated, "dispatch_gap_ms": gap,
            "bytes_per_launch": float(alg.mean()), "stored_bytes_per_launch": float(sto.mean()),
            "stored_bytes": algorithmic_bytes(blocks, npts, neq, fmt),
            "achieved": float(alg.sum() / t / 1e9), "achieved_stored": float(sto.sum() / t / 1e9)}


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


def measure_learned(ctx, u, r, npts, stream, dev, fused_ms):
    """Trains the branch predictor on pairs from a batch of small instances,
    then times, at the bench state, the reference's unfused steps with and
    without branch sorting (ComputeSigma on points grouped by predicted
    branch); the results of both are identical."""
    import numpy as np
    import torch

    from intpointsfem_amd import learned as L
    t0 = time.time()
    f_tr, b_tr, t_tr = L.training_set(n_instances=24, seed=1)
    mlp = L.new_mlp(seed=0)
    hist = L.train(mlp, f_tr, b_tr, t_tr, steps=3000, batch=4096, lr=3e-3, log_every=1000)
    train_s = time.time() - t0
    del f_tr, b_tr, t_tr
    feat, br_exact, tg = L.point_features(ctx, u, npts, dev)
    strain = torch.zeros(3 * npts, dtype=torch.float64, device=dev)
    ws = torch.zeros(3 * npts, dtype=torch.float64, device=dev)
    pred = torch.zeros(npts, dtype=torch.int32, device=dev)
    order = torch.zeros(npts, dtype=torch.int32, device=dev)
    ctx.set_update_mem(False)
    t = {
        "multiply_ms": event_ms(lambda: ctx.multiply(u, strain), stream),
        "features_ms": event_ms(lambda: ctx.point_features(u, feat, br_exact, tg), stream),
        "predict_ms": event_ms(lambda: mlp.predict(feat, branch=pred, n=npts), stream),
        "sort_ms": event_ms(lambda: ctx.branch_order(pred, order), stream),
        "compute_sigma_ms": event_ms(lambda: ctx.compute_sigma(strain, ws), stream),
        "compute_sigma_sorted_ms": event_ms(lambda: ctx.compute_sigma_ordered(strain, ws, order), stream),
        "multiply_transpose_ms": event_ms(lambda: ctx.multiply_transpose(ws, r), stream),
    }
    acc = float((pred == br_exact).float().mean().item())
    be = br_exact.cpu().numpy()
    div_nat = L.divergence(be, 63)
    div_sorted = L.divergence(be[order.cpu().numpy()], 64)
    plain = t["multiply_ms"] + t["compute_sigma_ms"] + t["multiply_transpose_ms"]
    learned = (t["multiply_ms"] + t["features_ms"] + t["predict_ms"] + t["sort_ms"] + t["compute_sigma_sorted_ms"]
               + t["multiply_transpose_ms"])
    log(f"[bench] learned: trained in {train_s:.1f}s (loss {hist[0][1]:.3f} -> {hist[-1][1]:.4f}), branch accuracy "
        f"{acc:.4f}; mixed-branch groups {div_nat:.3f} -> {div_sorted:.4f}; ComputeSigma {t['compute_sigma_ms']:.3f} "
        f"-> sorted {t['compute_sigma_sorted_ms']:.3f} ms; unfused residual plain {plain:.3f} ms, learned "
        f"{learned:.3f} ms, fused {fused_ms:.3f} ms")
    return dict({k: round(v, 4) for k, v in t.items()}, train_seconds=round(train_s, 1),
                branch_accuracy=round(acc, 5), mixed_branch_groups_natural=round(div_nat, 4),
                mixed_branch_groups_sorted=round(div_sorted, 5), unfused_residual_plain_ms=round(plain, 4),
                unfused_residual_learned_ms=round(learned, 4), fused_residual_ms=round(fused_ms, 4),
                learned_value=round(npts / learned / 1e3, 2))


def run_batch(args):
    """The 64 uneven wellbore instances (tests/golden/batch64_instances.json,
    drawn with seed 0) in one context: one launch per (color, shape) for the
    whole batch.  Same JSON line; cpu_baseline times the oracle on whole
    instances, in table order, until --cpu-seconds."""
    import numpy as np
    import torch

    import intpointsfem_amd as ipf
    from intpointsfem_amd import batch
    from intpointsfem_amd import wellbore as wb

    dev = torch.device("cuda:0")
    t0 = time.time()
    table = batch.draw_instances(64, seed=0)
    meshes = [batch.instance_mesh(i) for i in table]
    B = batch.Batch(meshes, device=0)
    ctx = B.ctx
    npts, neq = B.npts, B.neq
    log(f"[bench] batch of {B.n} instances: {len(B.row_sizes)} elements, {npts} points, neq {neq}; setup + upload "
        f"{time.time() - t0:.1f}s, {ctx.info()['device_bytes'] / 1e9:.2f} GB on device")
    stream = torch.cuda.current_stream(dev)
    ctx.set_stream(stream.cuda_stream)
    us = [wb.lame_displacement(m.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, inst["pw_final"], wb.SIGMA0),
                               E_=inst["E"], nu=inst["nu"]) for inst, m in zip(table, meshes)]
    u_host = np.concatenate(us)
    u = torch.as_tensor(u_host, dtype=torch.float64, device=dev)
    r = torch.zeros(neq, dtype=torch.float64, device=dev)
    ctx.set_update_mem(True)
    ctx.residual(u, r)
    ctx.set_update_mem(False)
    ctx.synchronize()
    _, _, mt, br = ctx.get_state()
    ctx.set_state(np.zeros(6 * npts))
    branch_counts = {ipf.BRANCH_NAMES[i]: int((br == i).sum()) for i in range(5)}
    log(f"[bench] branches: {branch_counts}")
    tune = None
    if not args.no_autotune:
        tune = ctx.autotune(u, reps=5)
        log(f"[bench] autotune: {', '.join(f'{k} {v:.4f} ms' for k, v in tune['residual'].items())} -> "
            f"{tune['selected'][0]}")
        tune = {"residual_ms": {k: round(v, 5) for k, v in tune["residual"].items()}, "selected": tune["selected"][0]}
    fmt0 = ctx.b_format()[0]
    nsettle = settle(ctx, u, r, args.settle_ms)
    total_ms = time_residuals(ctx, u, r, stream, args.steps, args.warmup, roi="timed residuals")
    ms_per_step = total_ms / args.steps
    value = npts * args.steps / (total_ms / 1e3) / 1e6
    log(f"[bench] {args.steps} batched residuals ({fmt0} B): {ms_per_step:.4f} ms, {value:.1f} Mint-pts/s")
    blocks = {"row_sizes": B.row_sizes, "col_sizes": B.col_sizes}
    nrep = max(3, min(args.steps, 20))
    roof = launch_roofline(ctx, u, r, nrep, blocks, npts, neq, fmt0, step_ms=ms_per_step)
    # the known-size k_copy that calibrates the PMC byte counters (see main)
    ctx.residual_tangent(u, torch.empty_like(r))
    ctx.synchronize()
    log(f"[bench] launches {', '.join(f'{x:.4f}' for x in roof['color_ms'])} ms; {roof['achieved']:.0f} GB/s on "
        f"plain FP64 bytes, {roof['achieved_stored']:.0f} GB/s on stored bytes")
    # HBM bytes per launch of each element order, from the committed rocprofv3
    # profile of this workload (PMC FETCH_SIZE + WRITE_SIZE, calibrated on k_copy)
    traffic, by_shape = None, None
    tfile = os.path.join(ROOT, "profiles", "k_residual_batch64_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tk = json.load(f)["kernels"]
        by_shape = {}
        for sh, nm in enumerate(("Q1", "Q2", "Q3")):
            kn = kernel_name(fmt0, sh)
            if kn in tk and "total_bytes" in tk[kn]:
                by_shape[nm] = {"traffic": int(tk[kn]["total_bytes"]), "valu_busy": tk[kn].get("valu_busy")}
        if len(by_shape) == 3 and len(roof["color_ms"]) == 3 * B.ncolor:
            # every color holds all three orders: the mean over the launches
            traffic = int(sum(v["traffic"] for v in by_shape.values()) / 3)
    roofline = {"bound": "hbm", "achieved": round(roof["achieved_stored"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(roof["achieved_stored"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_residual_fmt, one launch per (color, element order) over all instances",
                "algorithmic_bytes_per_point": round(roof["stored_bytes"] / npts, 1),
                "avg_launch_ms": round(roof["avg_launch_ms"], 5),
                "measured_traffic_rate": None if traffic is None else round(traffic / (roof["avg_launch_ms"] / 1e3) / 1e9, 1),
                "traffic_by_order": by_shape,
                "dense_fp64_bytes_rate": round(roof["achieved"], 1)}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        nthreads = max(1, min(16, os.cpu_count() or 1))
        done_pts, spent, ninst = 0, 0.0, 0
        for inst, m, ui in zip(table, meshes, us):
            sm = ipf.StructMatrix(m, device=-1)
            b = sm.export_blocks()
            ix = sm.export_indexes()
            s = sm.sizes()
            Bo = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"],
                          b["storage"])
            d = m.material()
            t = time.time()
            O.residual(Bo, ix["dof_indexes"], ix["color_indexes"], s["ncolor"], s["neq"], ix["weight"], d["lambda"],
                       d["mu"], d["phi"], d["psi"], d["cohesion"], ui, f_lin=ix["rhs_linear"], nthreads=nthreads)
            spent += time.time() - t
            done_pts += s["npts"]
            ninst += 1
            if spent >= args.cpu_seconds:
                break
        cpu = {"value": round(done_pts / spent / 1e6, 3), "unit": "Mint-pts/s", "cores": nthreads, "kind": "port",
               "sample": f"one residual of each of the first {ninst} instances ({done_pts} points) by the C oracle "
                         f"with {nthreads} OpenMP threads, {spent:.1f} s"}
        log(f"[bench] cpu oracle: {cpu['value']:.2f} Mint-pts/s ({ninst} instances)")
    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mint-pts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "plain FP64 batched residual, 64 uneven wellbore instances (seed-0 table)",
                      "b_format": fmt0, "autotune": tune, "settle": {"ms": args.settle_ms, "residuals": nsettle},
                      "instances": B.n, "elements": int(len(B.row_sizes)), "int_points": npts,
                      "neq": neq, "colors": B.ncolor, "branches": branch_counts,
                      "state": "u = each instance's linear-elastic thick-cylinder displacement, zero plastic strain"},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed back-to-back residuals before the warm-up steps, in ms of wall time")
    ap.add_argument("--mesh", default="800x800-q2",
                    help="800x800-q2 (default), 800x800-q3, 40x40-q2, or batch64 (64 uneven instances)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-formats", action="store_true", help="skip the compact / FP32 B format variants")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned branch-predictor measurement")
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the path runs on a single "
              "MI355X; multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    if args.mesh == "batch64":
        return run_batch(args)
    import numpy as np
    import torch

    import intpointsfem_amd as ipf
    from intpointsfem_amd import wellbore as wb

    dev = torch.device("cuda:0")
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
        # displacement: linear-elastic thick-cylinder field of the final load step
        u_host = wb.lame_displacement(mesh.nodes(), wb.normal_stress(wb.NSTEPS), scale=1.0)
        u = torch.as_tensor(u_host, dtype=torch.float64, device=dev)
        # branch mix of this state (one committed evaluation, then reset the state)
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
        # FP64 residual; tangent products in the chosen format (FP32: mixed Newton)
        solve_tf = args.solve_tangent_format or fmt_default[1]
        ctx.set_b_format(fmt_default[0], tangent=solve_tf)
        f_unit = sm.export_indexes()["rhs_linear"] / wb.normal_stress(wb.NSTEPS)
        u = torch.zeros(neq, dtype=torch.float64, device=dev)
        t1 = time.time()
        steps = []
        for k in range(1, wb.NSTEPS + 1):
            ctx.set_rhs_linear(f_unit * wb.normal_stress(k))
            st = ctx.newton(u, rtol=args.newton_rtol, maxit=60, cg_rtol=args.cg_rtol, cg_maxit=200000)
            steps.append(st)
            log(f"[bench] load step {k}: newton {int(st['iterations'])}, |r|/|r0| {st['rel_residual']:.2e}, "
                f"cg {int(st['cg_iterations'])}, {time.time() - t1:.1f}s")
        ctx.set_b_format(*fmt_default)
        u_host = u.cpu().numpy()
        _, _, mt, br = ctx.get_state()
        newton_info = {"solve_tangent_format": solve_tf, "load_steps": wb.NSTEPS, "newton_iterations": [int(s["iterations"]) for s in steps],
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
    nbytes = algorithmic_bytes(blocks, npts, neq)  # plain FP64 byte count: dense B
    roof = launch_roofline(ctx, u, r, nrep, blocks, npts, neq, fmt0, step_ms=ms_per_step)
    # one untimed ipf_residual_tangent: its k_copy (f_lin into r, neq doubles
    # read and written) is the known-size dispatch that calibrates the PMC
    # byte counters of a profiled run (the residual itself needs no copy)
    ctx.residual_tangent(u, torch.empty_like(r))
    ctx.synchronize()
    log(f"[bench] copy kernel {roof['copy_ms']:.4f} ms; color kernels "
        f"{', '.join(f'{x:.4f}' for x in roof['color_ms'])} ms; {nbytes / npts:.1f} algorithmic B/pt -> "
        f"{roof['achieved']:.0f} GB/s; {roof['stored_bytes'] / npts:.1f} stored B/pt -> {roof['achieved_stored']:.0f} GB/s")
    # HBM bytes per launch measured by rocprofv3 PMC (FETCH_SIZE, WRITE_SIZE calibrated on
    # k_copy) for this workload, committed with the profile it comes from
    kname = kernel_name(fmt0, shape)
    traffic = None
    valu_busy = None
    tfile = os.path.join(ROOT, "profiles", f"k_residual_{args.mesh}_traffic.json")
    # bytes per launch hardly depend on the state (same loads and stores at every point)
    if os.path.exists(tfile):
        with open(tfile) as f:
            tk = json.load(f)["kernels"]
        if kname in tk:
            traffic = int(tk[kname]["total_bytes"]) if "total_bytes" in tk[kname] else None
            valu_busy = tk[kname].get("valu_busy")
    roofline = {"bound": "hbm", "achieved": round(roof["achieved_stored"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(roof["achieved_stored"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": f"{kname} (one launch per element color)",
                "algorithmic_bytes_per_launch": int(roof["stored_bytes_per_launch"]),
                "algorithmic_bytes_per_point": round(roof["stored_bytes"] / npts, 1),
                "avg_launch_ms": round(roof["avg_launch_ms"], 5),
                "avg_launch_ms_isolated": round(roof["avg_launch_ms_isolated"], 5),
                "measured_traffic_rate": None if traffic is None else round(traffic / (roof["avg_launch_ms"] / 1e3) / 1e9, 1),
                # 4 cycles x VALU instructions / (1024 SIMDs x 2.4 GHz x duration), rocprofv3 PMC (profiles/)
                "valu_busy": None if valu_busy is None else round(valu_busy, 3),
                "dense_fp64_bytes_rate": round(roof["achieved"], 1)}

    # the other B formats (same state, same u): timing, roofline on their own
    # stored bytes, and the deviation from the FP64 residual
    variants = {}
    if not args.no_formats:
        ctx.residual(u, r)
        ctx.synchronize()
        r64 = r.cpu().numpy()
        fint = float(np.abs(r64 - sm.export_indexes()["rhs_linear"]).max())
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

    learned_info = None if args.no_learned else measure_learned(ctx, u, r, npts, stream, dev, ms_per_step)

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
        "b_format_variants": variants,
        "learned": learned_info,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
