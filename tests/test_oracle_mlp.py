"""The CPU reference of the learned predictor (oracle/mlp.py): float64
back-propagation against central finite differences of the loss, the
parameter count of the 7 -> 64 -> 64 -> 6 network, and the BF16-rounded
gradient (the one the GPU kernels mirror) against the exact one."""
import numpy as np

from oracle import mlp as M


def _data(rng, n):
    feat = rng.standard_normal((n, 8)).astype(np.float32)
    label = rng.integers(0, 5, size=n)
    target = rng.standard_normal(n)
    return feat, label, target


def _params(rng, scale=0.3):
    p = rng.standard_normal(M.NPAR) * scale
    p[~M.trainable_mask()] = 0.0
    return p


def test_parameter_count():
    assert M.trainable_mask().sum() == 7 * 64 + 64 + 64 * 64 + 64 + 64 * 6 + 6 == 5062


def _min_preactivation(p, feat):
    W1, b1, W2, b2, W3, b3 = M.unpack(np.asarray(p, dtype=np.float64))
    Z1 = M._inputs(feat).astype(np.float64) @ W1 + b1
    Z2 = np.fmax(Z1, 0) @ W2 + b2
    return min(np.abs(Z1).min(), np.abs(Z2).min())


def test_backprop_matches_finite_differences():
    """Central differences of the float64 loss against gradient64.  ReLU has
    a kink at 0: the data and parameters are drawn (first seed from 11 on)
    so that every pre-activation is at least 1e-4 away from 0, while a step
    of h = 1e-7 in one parameter moves none by more than ~1e-5, so the
    differences never straddle a kink and the loss is linear in each
    activation's pattern between them."""
    for seed in range(11, 60):
        rng = np.random.default_rng(seed)
        feat, label, target = _data(rng, 40)
        p = _params(rng)
        if _min_preactivation(p, feat) >= 1e-4:
            break
    else:
        raise AssertionError("no seed keeps the pre-activations away from the ReLU kink")
    _, g = M.gradient64(p, feat, label, target)
    idx = np.concatenate([rng.choice(np.nonzero(M.trainable_mask()[a:b])[0] + a, 6, replace=False)
                          for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                                       (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.OFF_B3 + 6))])
    h = 1e-7
    for i in idx:
        e = np.zeros(M.NPAR)
        e[i] = h
        fd = (M.loss64(p + e, feat, label, target) - M.loss64(p - e, feat, label, target)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])
    assert (g[~M.trainable_mask()] == 0).all()


def test_bf16_gradient_close_to_exact():
    rng = np.random.default_rng(12)
    feat, label, target = _data(rng, 512)
    p = _params(rng, 0.2)
    l64, g64 = M.gradient64(p, feat, label, target)
    l16, g16 = M.gradient(p, feat, label, target)
    assert abs(l16 - l64) <= 2e-2 * abs(l64)
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159], dtype=np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to the even mantissa (1)
    assert M.bf16(x)[0] == 1.0 and M.bf16(x)[1] == 1.0 and M.bf16(x)[2] == 1.0078125
