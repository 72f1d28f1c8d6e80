"""Cubic-spline kernel tables (reference: wp_sph/solver/sph_funcs.py:43-89).

The reference tabulates W(r) and (grad W)/r on ``linspace(0, h, N)`` in
float64 and casts to float32 (sph_funcs.py:51-86). The same closed forms are
evaluated here with the same numpy operations, so the float32 tables are
identical for equal (h, N). The device kernels interpolate them exactly as
``spline_W`` / ``grad_spline_W`` do (sph_funcs.py:92-118).
"""
from __future__ import annotations

import math

import numpy as np

#: table sizes of the reference (sph_funcs.py:43-44)
W_SIZE = 8192
GRAD_W_SIZE = 65536


def sigma_of(h: float) -> float:
    """SIGMA = 8 / (pi h^3) (wp_sph/main.py:60)."""
    return 8.0 / (math.pi * h ** 3)


def w_table(h: float, size: int) -> np.ndarray:
    """W(r) = sigma [6 (q^3 - q^2) + 1] for q <= 1/2, 2 sigma (1 - q)^3 for
    1/2 < q < 1, else 0, with q = r / h (sph_funcs.py:56-71)."""
    sigma = sigma_of(h)
    h_inv = 1.0 / h
    r = np.linspace(0, h, size, dtype=np.float64)
    q = r * h_inv
    out = np.zeros_like(r)
    low = q <= 0.5
    high = (q > 0.5) & (q < 1.0)
    out[low] = (6.0 * (q[low] ** 3 - q[low] ** 2) + 1.0) * sigma
    out[high] = 2.0 * np.power(1.0 - q[high], 3) * sigma
    return out.astype(np.float32)


def grad_table(h: float, size: int) -> np.ndarray:
    """(-dW/dr) / r: (12 - 18 q) sigma / h^2 for q <= 1/2,
    6 (1 - q)^2 sigma / (h r) for 1/2 < q < 1, 0 at r <= 1e-10
    (sph_funcs.py:73-82)."""
    sigma = sigma_of(h)
    h_inv = 1.0 / h
    h_inv2 = 1.0 / (h * h)
    r = np.linspace(0, h, size, dtype=np.float64)
    q = r * h_inv
    out = np.zeros_like(r)
    nz = r > 1e-10
    low = nz & (q <= 0.5)
    high = nz & (q > 0.5) & (q < 1.0)
    out[low] = (12.0 - 18.0 * q[low]) * sigma * h_inv2
    out[high] = 6.0 * np.power(1.0 - q[high], 2) * sigma * h_inv / r[high]
    return out.astype(np.float32)


def make_tables(h: float, w_size: int = W_SIZE, g_size: int = GRAD_W_SIZE):
    """Both tables as float32 arrays (precompute_tables, sph_funcs.py:51-86)."""
    return w_table(h, w_size), grad_table(h, g_size)


def spline_w_lookup(r, table: np.ndarray, h: float) -> np.ndarray:
    """Vectorised float32 ``spline_W`` (sph_funcs.py:92-102), for host-side
    features and tests."""
    r = np.asarray(r, dtype=np.float32)
    n = table.shape[0]
    hf = np.float32(h)
    scale = np.float32(n - 1) * np.float32(1.0 / h)
    idx = r * scale
    lo = np.minimum(idx.astype(np.int32), n - 2)
    alpha = idx - lo.astype(np.float32)
    val = (np.float32(1.0) - alpha) * table[lo] + alpha * table[lo + 1]
    return np.where(r >= hf, np.float32(0.0), val).astype(np.float32)
