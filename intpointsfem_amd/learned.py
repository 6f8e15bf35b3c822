"""Learned return-branch predictor: training pairs from the exact return map,
training on the GPU (the library's own BF16-MFMA forward / backward and Adam
kernels), and the branch-sorted ComputeSigma that uses it.

The predictor only orders the points: ``compute_sigma_sorted`` runs the
exact return map on every point (``ipf_compute_sigma_ordered``), so its
results are those of ``ipf_compute_sigma`` whatever the prediction.
"""
import numpy as np

from . import MLP
from . import batch as _batch
from . import wellbore as wb


def point_features(ctx, u, npts, device):
    """(feat [npts, 8] float32, branch [npts] int32, target [npts] float32)
    device tensors: the exact return map's pairs at the trial state of u."""
    import torch
    feat = torch.zeros((npts, 8), dtype=torch.float32, device=device)
    br = torch.zeros(npts, dtype=torch.int32, device=device)
    tg = torch.zeros(npts, dtype=torch.float32, device=device)
    ctx.point_features(u, feat, br, tg)
    return feat, br, tg


def training_set(n_instances=24, N=(12, 40), scales=(0.5, 1.0, 1.5, 2.0, 3.0, 5.0), seed=1, device="cuda:0"):
    """Pairs from a batch of small instances drawn like the 64-instance
    configuration (materials and pressures from batch.draw_instances, mesh
    sizes N in the given range) at several multiples of each instance's
    linear-elastic displacement.  Returns (feat, branch, target) tensors."""
    import torch
    rng = np.random.default_rng(seed)
    table = _batch.draw_instances(n_instances, seed=seed)
    for inst in table:
        inst["N"] = int(rng.integers(N[0], N[1], endpoint=True))
    meshes = [_batch.instance_mesh(i) for i in table]
    B = _batch.Batch(meshes, device=0)
    B.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    us = [wb.lame_displacement(m.nodes(), wb.normal_stress(wb.NSTEPS, wb.NSTEPS, i["pw_final"], wb.SIGMA0),
                               E_=i["E"], nu=i["nu"]) for i, m in zip(table, meshes)]
    u0 = np.concatenate(us)
    parts = []
    for s in scales:
        u = torch.as_tensor(u0 * s, dtype=torch.float64, device=device)
        parts.append(point_features(B.ctx, u, B.npts, device))
    B.ctx.synchronize()
    return tuple(torch.cat([p[k] for p in parts]) for k in range(3))


def new_mlp(seed=0, device=0):
    """An MLP whose kernels run on torch's current stream (ordered with the
    torch operations that prepare its minibatches)."""
    import torch
    m = MLP(device=device, seed=seed)
    m.set_stream(torch.cuda.current_stream().cuda_stream)
    return m


def train(mlp, feat, branch, target, steps=2000, batch=4096, lr=3e-3, seed=0, log_every=0):
    """Adam on random minibatches (indexes drawn on the host from `seed`,
    gathered on the device), learning rate on a cosine decay to 10 %;
    returns (step, loss) every `log_every` steps."""
    import torch
    n = branch.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(feat.device)
        f = feat.index_select(0, idx).contiguous()
        b = branch.index_select(0, idx).contiguous()
        t = target.index_select(0, idx).contiguous()
        want = log_every > 0 and (it % log_every == 0 or it == steps - 1)
        # cosine decay to 10 % of lr
        lr_t = lr * (0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * it / steps)))
        loss = mlp.train_step(f, b, t, lr_t, want_loss=want)
        if want:
            hist.append((it, loss))
    return hist


def predict(mlp, feat):
    import torch
    n = feat.shape[0]
    br = torch.zeros(n, dtype=torch.int32, device=feat.device)
    mlp.predict(feat, branch=br, n=n)
    return br


def compute_sigma_sorted(ctx, mlp, strain, wsigma, feat, order):
    """ComputeSigma with the points visited grouped by predicted branch:
    predict, stable counting sort, exact return map in that order."""
    br = predict(mlp, feat)
    ctx.branch_order(br, order)
    ctx.compute_sigma_ordered(strain, wsigma, order)
    return br


def divergence(branch_in_visit_order, width=64):
    """Fraction of `width`-point groups (in visiting order) with more than
    one branch: the wavefronts whose lanes take different returns."""
    b = np.asarray(branch_in_visit_order)
    n = len(b) // width * width
    g = b[:n].reshape(-1, width)
    return float((g != g[:, :1]).any(axis=1).mean())


__all__ = ["MLP", "new_mlp", "point_features", "training_set", "train", "predict", "compute_sigma_sorted", "divergence"]
