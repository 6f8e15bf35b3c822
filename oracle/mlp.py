"""CPU reference of the learned return-branch predictor (test infrastructure:
the checker of the GPU kernels in intpointsfem_amd/csrc/ipf_learned.hip,
never part of the product path).

7 -> 64 -> 64 -> 6 MLP with ReLU activations.
`forward` / `gradient` round the matrix-core operands to BF16 (round to
nearest even) where the GPU does -- inputs, weights, activations and the
back-propagated dL/dZ -- and compute in float32; `loss64` / `gradient64`
are the same network in float64 without rounding, for finite-difference
checks of the back-propagation formulas.

Parameter vector layout (padded entries zero): W1 [16 x 64], b1 [64],
W2 [64 x 64], b2 [64], W3 [64 x 16], b3 [16]; row-major (input, output).
"""
import numpy as np

IN, INP, H, NBR, OUT, OUTP = 7, 16, 64, 5, 6, 16
REG = 0.1
OFF_W1 = 0
OFF_B1 = OFF_W1 + INP * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * OUTP
NPAR = OFF_B3 + OUTP


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    return (p[OFF_W1:OFF_B1].reshape(INP, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(H, OUTP), p[OFF_B3:NPAR])


def trainable_mask():
    m = np.ones(NPAR, dtype=bool)
    W1 = m[OFF_W1:OFF_B1].reshape(INP, H)
    W1[IN:, :] = False
    W3 = m[OFF_W3:OFF_B3].reshape(H, OUTP)
    W3[:, OUT:] = False
    m[OFF_B3 + OUT:NPAR] = False
    return m


def _relu(x):
    """max(x, 0) (the GPU takes it on the bit pattern: the same for every
    non-NaN value, -0 to +0)."""
    return np.fmax(x, 0)


def _relu_grad(x):
    """1 above 0, else 0 (the GPU kernels' convention at 0)."""
    return (x > 0).astype(x.dtype)


def _inputs(feat):
    X = np.zeros((len(feat), INP), dtype=np.float32)
    X[:, :IN] = np.asarray(feat, dtype=np.float32).reshape(len(feat), -1)[:, :IN]
    return X


def forward(p, feat):
    """(Z1, H1, Z2, H2, Z3, X) with the GPU's BF16 operand rounding."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(a, dtype=np.float32) for a in unpack(p)]
    X = bf16(_inputs(feat))
    Z1 = X @ bf16(W1) + b1
    H1 = bf16(_relu(Z1))
    Z2 = H1 @ bf16(W2) + b2
    H2 = bf16(_relu(Z2))
    Z3 = H2 @ bf16(W3) + b3
    return Z1, H1, Z2, H2, Z3, X


def _loss_grad_out(Z3, label, target, n):
    z = Z3[:, :NBR].astype(np.float64)
    zmax = z.max(axis=1, keepdims=True)
    e = np.exp(z - zmax)
    se = e.sum(axis=1, keepdims=True)
    lp = z - zmax - np.log(se)
    idx = np.arange(len(label))
    diff = Z3[:, NBR].astype(np.float64) - target
    loss = (-lp[idx, label] + REG * diff * diff).sum() / n
    D = np.zeros((len(label), OUTP))
    D[:, :NBR] = e / se
    D[idx, label] -= 1.0
    D[:, NBR] = 2 * REG * diff
    return loss, D / n


def gradient(p, feat, label, target):
    """(loss, gradient [NPAR]) of the mean loss, mirroring the GPU kernels."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(a, dtype=np.float32) for a in unpack(p)]
    n = len(label)
    Z1, H1, Z2, H2, Z3, X = forward(p, feat)
    loss, D = _loss_grad_out(Z3, np.asarray(label), np.asarray(target, dtype=np.float64), n)
    D3 = bf16(D.astype(np.float32))
    g = np.zeros(NPAR, dtype=np.float64)
    g[OFF_W3:OFF_B3] = (H2.T @ D3).reshape(-1)
    g[OFF_B3:NPAR] = D3.sum(axis=0)
    D2 = bf16((D3 @ bf16(W3).T) * _relu_grad(Z2))
    g[OFF_W2:OFF_B2] = (H1.T @ D2).reshape(-1)
    g[OFF_B2:OFF_W3] = D2.sum(axis=0)
    D1 = bf16((D2 @ bf16(W2).T) * _relu_grad(Z1))
    g[OFF_W1:OFF_B1] = (X.T @ D1).reshape(-1)
    g[OFF_B1:OFF_W2] = D1.sum(axis=0)
    g[~trainable_mask()] = 0.0
    return loss, g


def loss64(p, feat, label, target):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, dtype=np.float64))
    X = _inputs(feat).astype(np.float64)
    Z3 = _relu(_relu(X @ W1 + b1) @ W2 + b2) @ W3 + b3
    return _loss_grad_out(Z3, np.asarray(label), np.asarray(target, dtype=np.float64), len(label))[0]


def gradient64(p, feat, label, target):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, dtype=np.float64))
    X = _inputs(feat).astype(np.float64)
    Z1 = X @ W1 + b1
    H1 = _relu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = _relu(Z2)
    Z3 = H2 @ W3 + b3
    loss, D3 = _loss_grad_out(Z3, np.asarray(label), np.asarray(target, dtype=np.float64), len(label))
    g = np.zeros(NPAR)
    g[OFF_W3:OFF_B3] = (H2.T @ D3).reshape(-1)
    g[OFF_B3:NPAR] = D3.sum(axis=0)
    D2 = (D3 @ W3.T) * _relu_grad(Z2)
    g[OFF_W2:OFF_B2] = (H1.T @ D2).reshape(-1)
    g[OFF_B2:OFF_W3] = D2.sum(axis=0)
    D1 = (D2 @ W2.T) * _relu_grad(Z1)
    g[OFF_W1:OFF_B1] = (X.T @ D1).reshape(-1)
    g[OFF_B1:OFF_W2] = D1.sum(axis=0)
    g[~trainable_mask()] = 0.0
    return loss, g
