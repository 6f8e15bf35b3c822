"""The learned return-branch predictor on the GPU.

Bars:
  * training pairs (ipf_point_features): branch ids bit-identical to the
    oracle's return map, principal stresses and alpha increments equal to
    the oracle's values rounded to float32;
  * MLP forward (BF16 matrix-core operands, FP32 accumulation): outputs within
    1e-3 of the CPU mirror (oracle/mlp.py, same BF16 roundings), argmax equal
    on >= 99.9 % of samples;
  * gradient of the mean loss: within 1e-2 (relative 2-norm, per parameter
    block) of the CPU mirror, loss within 1e-4 relative; the mirror's
    back-propagation is pinned by finite differences (tests/test_oracle_mlp.py);
  * training on pairs from a batch of instances: held-out branch accuracy
    >= 99 % and far above the majority-class rate;
  * branch-sorted ComputeSigma: identical results (stresses, flags, plastic
    strain) to the unsorted one; the order groups points by predicted branch;
    with corrupted (NaN) weights the order is still a permutation and the
    results are still identical.
"""
import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import learned
from intpointsfem_amd import wellbore as wb
from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(mesh="40x40-q2"):
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    b = sm.export_blocks()
    ix = sm.export_indexes()
    B = O.Blocks(b["row_sizes"], b["col_sizes"], b["matrix_position"], b["row_first"], b["col_first"], b["storage"])
    return m, sm, sm.context(), B, ix


def test_point_features_match_oracle(gpu):
    import torch
    m, sm, ctx, B, ix = _setup()
    npts = sm.sizes()["npts"]
    u_h = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0)
    feat, br, tg = learned.point_features(ctx, torch.as_tensor(u_h, device=gpu), npts, gpu)
    ctx.synchronize()
    mat = m.material()
    strain = O.multiply(B, ix["dof_indexes"], u_h)
    _, ep, al, mt, br_o = O.compute_sigma(strain, ix["weight"], mat["lambda"], mat["mu"], mat["phi"], mat["psi"],
                                          mat["cohesion"], update_mem=True)
    assert np.array_equal(br.cpu().numpy(), br_o)
    assert len(np.unique(br_o)) >= 2
    assert np.array_equal(tg.cpu().numpy(), (al * 1e3).astype(np.float32))
    # trial principal stresses of the first points, from the oracle's decomposition
    lam, mu = mat["lambda"], mat["mu"]
    c = (2.0 * mat["cohesion"] * np.cos(mat["phi"])) / (2.0 * np.cos(mat["phi"]))
    f = feat.cpu().numpy()
    for o in range(0, npts, 997):
        exx, gxy, eyy = strain[3 * o:3 * o + 3]
        s6 = np.array([(lam + 2 * mu) * exx + lam * eyy + lam * 0.0, mu * gxy, mu * 0.0,
                       lam * exx + (lam + 2 * mu) * eyy + lam * 0.0, mu * 0.0, lam * exx + lam * eyy + (lam + 2 * mu) * 0.0])
        ev, _, _ = O.spectral_decomposition(s6)
        np.testing.assert_allclose(f[o, :3], (ev / c).astype(np.float32), rtol=1e-6, atol=1e-6)
    assert np.allclose(f[:, 4], mat["phi"]) and np.allclose(f[:, 5], mat["psi"])


def _random_samples(gpu, n=3000, seed=5):
    import torch
    rng = np.random.default_rng(seed)
    feat = np.zeros((n, 8), dtype=np.float32)
    feat[:, :3] = np.sort(rng.normal(0, 3, (n, 3)), axis=1)[:, ::-1]
    feat[:, 3] = rng.uniform(0, 3, n)
    feat[:, 4:6] = rng.uniform(0.35, 0.7, (n, 2))
    feat[:, 6] = rng.exponential(1.0, n)
    label = rng.integers(0, 5, n).astype(np.int32)
    target = rng.normal(0, 1, n).astype(np.float32)
    return (feat, label, target), tuple(torch.as_tensor(a, device=gpu) for a in (feat, label, target))


def test_forward_and_gradient_match_cpu_mirror(gpu):
    import torch
    (feat, label, target), (fd, ld, td) = _random_samples(gpu)
    mlp = learned.new_mlp(seed=3)
    p = mlp.params().astype(np.float64)
    n = len(label)
    out = torch.zeros((n, 6), dtype=torch.float32, device=gpu)
    br = torch.zeros(n, dtype=torch.int32, device=gpu)
    mlp.predict(fd, branch=br, out=out, n=n)
    torch.cuda.synchronize()
    Z3 = M.forward(p, feat)[4][:, :6]
    og = out.cpu().numpy()
    assert np.abs(og - Z3).max() <= 1e-3 * max(1.0, np.abs(Z3).max())
    assert (br.cpu().numpy() == Z3[:, :5].argmax(axis=1)).mean() >= 0.999
    loss_g, g = mlp.gradient(fd, ld, td)
    loss_c, gc = M.gradient(p, feat, label, target)
    print(f"loss gpu {loss_g:.6f} cpu {loss_c:.6f}")
    assert abs(loss_g - loss_c) <= 1e-4 * abs(loss_c)
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPAR)):
        err = np.linalg.norm(g[a:b] - gc[a:b]) / np.linalg.norm(gc[a:b])
        print(f"block [{a}, {b}): relative gradient error {err:.2e}")
        assert err <= 1e-2
    assert (g[~M.trainable_mask()] == 0).all()


def test_training_reaches_high_branch_accuracy(gpu):
    import torch
    feat, br, tg = learned.training_set(n_instances=24, seed=1)
    n = br.shape[0]
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(0)).to(gpu)
    ntr = int(0.8 * n)
    tr, te = perm[:ntr], perm[ntr:]
    f_tr, b_tr, t_tr = (x.index_select(0, tr).contiguous() for x in (feat, br, tg))
    f_te, b_te = (x.index_select(0, te).contiguous() for x in (feat, br))
    mlp = learned.new_mlp(seed=0)
    hist = learned.train(mlp, f_tr, b_tr, t_tr, steps=3000, batch=4096, lr=3e-3, log_every=500)
    pred = learned.predict(mlp, f_te)
    torch.cuda.synchronize()
    y, yp = b_te.cpu().numpy(), pred.cpu().numpy()
    acc = float((y == yp).mean())
    counts = np.bincount(y, minlength=5)
    major = counts.max() / counts.sum()
    recall = [float((yp[y == k] == k).mean()) if counts[k] else float("nan") for k in range(5)]
    print(f"{n} pairs, class counts {counts.tolist()}; loss {hist[0][1]:.3f} -> {hist[-1][1]:.4f}; "
          f"held-out accuracy {acc:.4f} (majority class {major:.3f}); recall per branch "
          f"{[round(r, 3) for r in recall]}")
    assert acc >= 0.99 and acc > major + 0.5 * (1 - major)


def test_branch_sorted_compute_sigma_is_exact(gpu):
    import torch
    m, sm, ctx, B, ix = _setup()
    s = sm.sizes()
    npts = s["npts"]
    u_h = wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0)
    u = torch.as_tensor(u_h, device=gpu)
    feat, br_exact, _ = learned.point_features(ctx, u, npts, gpu)
    strain = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ctx.multiply(u, strain)
    ws0 = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ctx.set_update_mem(True)
    ctx.compute_sigma(strain, ws0)
    ctx.synchronize()
    st0 = ctx.get_state()
    ctx.set_state(np.zeros(6 * npts))
    # an untrained network gives some order; the results must not depend on it
    mlp = learned.new_mlp(seed=7)
    order = torch.zeros(npts, dtype=torch.int32, device=gpu)
    ws1 = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    pred = learned.compute_sigma_sorted(ctx, mlp, strain, ws1, feat, order)
    ctx.set_update_mem(False)
    ctx.synchronize()
    st1 = ctx.get_state()
    assert (ws1.cpu().numpy() == ws0.cpu().numpy()).all()
    for a, b in zip(st0, st1):
        assert np.array_equal(a, b)
    o = order.cpu().numpy()
    p = pred.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(npts))
    assert (np.diff(p[o]) >= 0).all()
    # exact branches as the key: the sorted visiting order has at most 4 mixed wavefronts
    ctx.branch_order(br_exact, order)
    ctx.synchronize()
    be = br_exact.cpu().numpy()
    d_sorted = learned.divergence(be[order.cpu().numpy()])
    d_natural = learned.divergence(be)
    print(f"mixed-branch wavefronts: natural order {d_natural:.4f}, sorted by exact branch {d_sorted:.4f}")
    assert d_sorted * (npts // 64) <= 4


def test_corrupted_weights_keep_results_exact(gpu):
    """NaN weights: the predictions are garbage, the order is still a
    permutation and ComputeSigma's results are unchanged."""
    import torch
    m, sm, ctx, B, ix = _setup()
    npts = sm.sizes()["npts"]
    u = torch.as_tensor(wb.lame_displacement(m.nodes(), wb.normal_stress(10), scale=6.0), device=gpu)
    feat, _, _ = learned.point_features(ctx, u, npts, gpu)
    strain = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ctx.multiply(u, strain)
    ws0 = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    ctx.compute_sigma(strain, ws0)
    ctx.synchronize()
    mlp = learned.new_mlp(seed=1)
    p = mlp.params()
    p[M.OFF_W2:M.OFF_B2] = np.nan
    mlp.set_params(p)
    order = torch.zeros(npts, dtype=torch.int32, device=gpu)
    ws1 = torch.zeros(3 * npts, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    learned.compute_sigma_sorted(ctx, mlp, strain, ws1, feat, order)
    ctx.synchronize()
    assert np.array_equal(np.sort(order.cpu().numpy()), np.arange(npts))
    assert (ws1.cpu().numpy() == ws0.cpu().numpy()).all()
