"""Ablation of the learned + mixed batched residual on the 64 uneven
instances (seed-0 table), at a committed plastic state (u1 = 2x each
instance's linear-elastic displacement committed, residuals at u2 = 3x):

  plain                        colored assembly, exact return map, index order
  + certified elastic          points proven elastic skip the spectral decomposition
  + predicted order            the predictor (trained on this state's pairs)
                               orders each color run / patch so predicted-elastic
                               elements share tiles
  patch assembly               each variant again with the patch assembly

Bars: every variant within 1e-10 of max|f_int| of the plain residual (the
plain one is pinned to the oracle at ~1e-16 by test_gpu_production_oracle);
the colored certified residual bit-identical in index and predicted order
(within a color every DoF has one element); the best learned + mixed
variant at least 1.05x faster than plain (measured ~1.2x: the certified
path removes the return map's instructions from two thirds of the points,
the colored gather / scatter traffic remains).  Times are printed.
"""
import numpy as np
import pytest

from intpointsfem_amd import batch as bt
from intpointsfem_amd import learned
from intpointsfem_amd import wellbore as wb

pytestmark = pytest.mark.gpu


def _time(ctx, u, r, reps=30):
    import torch
    s = torch.cuda.current_stream()
    for _ in range(5):
        ctx.residual(u, r)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        ctx.residual(u, r)
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def test_learned_mixed_batched_ablation(gpu):
    import torch
    table = bt.draw_instances(64, seed=0)
    meshes = [bt.instance_mesh(i) for i in table]
    B = bt.Batch(meshes, device=0)
    ctx = B.ctx
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    us = [wb.lame_displacement(m.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, i["pw_final"], wb.SIGMA0),
                               E_=i["E"], nu=i["nu"]) for i, m in zip(table, meshes)]
    u1 = torch.as_tensor(np.concatenate([2 * x for x in us]), device=gpu)
    u2 = torch.as_tensor(np.concatenate([3 * x for x in us]), device=gpu)
    r = torch.zeros(B.neq, dtype=torch.float64, device=gpu)
    ctx.set_update_mem(True)
    ctx.residual(u1, r)
    ctx.set_update_mem(False)
    ctx.synchronize()
    times, res = {}, {}

    def run(label, asm, mode):
        ctx.set_residual_assembly(asm)
        ctx.set_residual_mode(mode)
        times[label] = _time(ctx, u2, r)
        ctx.residual(u2, r)
        ctx.synchronize()
        res[label] = r.cpu().numpy().copy()

    run("plain", "colored", "exact")
    run("certified", "colored", "certified-elastic")
    run("certified, patch", "patch", "certified-elastic")
    feat, br, tg = learned.point_features(ctx, u2, B.npts, gpu)
    mlp = learned.new_mlp(seed=0)
    learned.train(mlp, feat, br, tg, steps=2000, batch=4096, lr=3e-3, seed=0)
    pred = learned.predict(mlp, feat)
    torch.cuda.synchronize()
    acc = float((pred == br).float().mean())
    nip = (np.asarray(B.row_sizes) // 3).astype(np.int64)
    first = np.concatenate([[0], np.cumsum(nip)[:-1]])
    keys = (np.maximum.reduceat((pred.cpu().numpy() != 0).astype(np.int32), first) > 0).astype(np.int32)
    ctx.set_residual_assembly("colored")
    ctx.set_residual_mode("exact")
    B.relayout(keys)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    run("certified, predicted order", "colored", "certified-elastic")
    run("certified, predicted order, patch", "patch", "certified-elastic")
    ctx.set_residual_assembly("colored")
    ctx.set_residual_mode("exact")
    flin = B.rhs_linear
    scale = np.abs(res["plain"] - flin).max()
    print(f"predictor accuracy at u2 {acc:.4f}; elements keyed plastic {int(keys.sum())} of {len(keys)}")
    for k, t in times.items():
        dev = np.abs(res[k] - res["plain"]).max() / scale
        print(f"{k:36s} {t:.4f} ms  {times['plain'] / t:.3f}x plain  deviation {dev:.2e}", flush=True)
        assert dev <= 1e-10, k
    assert np.array_equal(res["certified"], res["certified, predicted order"])
    best = min(t for k, t in times.items() if k != "plain")
    assert times["plain"] / best >= 1.05
