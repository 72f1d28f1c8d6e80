"""CPU restatement of the per-particle pressure predictor (8 -> 64 -> 64 -> 1,
SiLU) in float64 numpy.

TEST INFRASTRUCTURE ONLY: the oracle of the BF16 matrix-core kernels in
warpsph_amd/csrc/mlp.hip. Only tests/ import this module.

The network has no counterpart in the reference (it replaces the reference's
p = 0 start of the Jacobi loop, wp_sph/solver/IISPH.py:166-174, :1009-1013).
Parameter vector layout (float32, 4801 values): W1[64][8], b1[64],
W2[64][64], b2[64], W3[64], b3.

``forward(..., bf16=True)`` models what the matrix cores see: the weights and
every matrix operand (inputs, hidden activations) are rounded to bfloat16
(round to nearest even); products and sums are exact here (float64), the
biases stay float32. ``grad`` is the gradient of 0.5 * mean (y - t)^2 with the
rounding treated as the identity (straight-through), which is what the
device backward pass computes. With ``bf16=False`` the model is smooth and
``grad`` is checked against central finite differences.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_H * N_IN
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H
N_PARAM = OFF_B3 + 1


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and return them as
    float64. Finite inputs only."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_H, N_IN), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(N_H, N_H), p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3],
            p[OFF_B3])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _r(a, on):
    """bf16 operand (from its float32 value) or the exact float64 value."""
    return bf16(np.asarray(a, dtype=np.float32)) if on else np.asarray(a, dtype=np.float64)


def forward(p, x, bf16: bool = True, keep: bool = False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xa = _r(x, bf16)
    z1 = xa @ _r(W1, bf16).T + b1
    h1 = silu(z1)
    h1a = _r(h1, bf16)
    W2a = _r(W2, bf16)
    z2 = h1a @ W2a.T + b2
    h2 = silu(z2)
    h2a = _r(h2, bf16)
    w3a = _r(W3, bf16)
    y = h2a @ w3a + b3
    if keep:
        return y, dict(xa=xa, z1=z1, h1a=h1a, z2=z2, h2a=h2a, W2a=W2a, w3a=w3a)
    return y


def loss(p, x, t, bf16: bool = True) -> float:
    y = forward(p, x, bf16)
    return 0.5 * float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def grad(p, x, t, bf16: bool = True):
    """(loss, gradient vector) of 0.5 mean (y - t)^2."""
    y, c = forward(p, x, bf16, keep=True)
    n = y.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    dy = e / n
    g = np.zeros(N_PARAM)
    g[OFF_W3:OFF_B3] = dy @ c["h2a"]
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, c["w3a"]) * dsilu(c["z2"])
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ c["h1a"]).ravel()
    dz1 = (dz2 @ c["W2a"]) * dsilu(c["z1"])
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (dz1.T @ c["xa"]).ravel()
    return 0.5 * float(np.mean(e * e)), g


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, algorithm 1) in float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
