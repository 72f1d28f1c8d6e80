"""CPU restatement of the learned start's feature pass and start value.

TEST INFRASTRUCTURE ONLY: the oracle of warpsph_amd/csrc/learn.hip (and the
feature seeds written by k_assemble). Only tests/ import this module. The
learned start has no counterpart in the reference, which starts every solve
from p = 0 (wp_sph/solver/IISPH.py:166-174).

Features of particle i, from the fields of one step (original order):
  0  (rho_adv / rho0 - 1) * 100
  1  a_ii / a_ref,  a_ref = -dt^2 m^2 / rho0^2 * sum_lattice |grad W|^2
  2  fluid neighbour count / 32
  3  carried start pressure / P_ref,  P_ref = rho0 * max(|g|, 1) * 0.1
  4-7 sum_j W_ij f_j / sum_j W_ij over i and its fluid neighbours (W_ii = W(0))
Columns 0-3 are clamped to +-64 before the means are taken. Computed in
float64 from float32 inputs.
"""
from __future__ import annotations

import math

import numpy as np


FEAT_CLAMP = 64.0
WEAK_COUPLING = 0.05


def lattice_grad_sq_sum(h: float, d: float) -> float:
    """sum over lattice points x = d (a, b, c), 0 < |x| < h, of |grad W(x)|^2
    for the cubic spline W = sigma [6 (q^3 - q^2) + 1] (q <= 1/2),
    2 sigma (1 - q)^3 (q < 1), sigma = 8 / (pi h^3)."""
    sigma = 8.0 / (math.pi * h ** 3)
    R = int(h / d) + 1
    a = np.arange(-R, R + 1)
    A, B, C = np.meshgrid(a, a, a, indexing="ij")
    r = d * np.sqrt((A * A + B * B + C * C).astype(np.float64)).ravel()
    r = r[(r > 0) & (r < h)]
    q = r / h
    # |dW/dr| from the derivative of the two polynomial pieces
    g = np.where(q <= 0.5, sigma * (18 * q * q - 12 * q) / h * -1.0,
                 sigma * 6 * (1 - q) ** 2 / h)
    return float(np.sum(g * g))


def a_ref(cfg, dt: float) -> float:
    return -dt * dt * cfg.mass ** 2 / cfg.rho0 ** 2 * lattice_grad_sq_sum(cfg.h, cfg.spacing)


def p_ref(cfg) -> float:
    return float(np.float32(cfg.rho0 * max(abs(cfg.gravity), 1.0) * 0.1))


def arbiter_error(p0, pf, a_ii, cfg, dt) -> float:
    """sqrt(mean((w (p0 - pf))^2)), w = min(|a_ii / a_ref|, 1)."""
    w = np.minimum(np.abs(np.asarray(a_ii, np.float64) / a_ref(cfg, dt)), 1.0)
    e = w * (np.asarray(p0, np.float64) - np.asarray(pf, np.float64))
    return float(np.sqrt(np.mean(e * e)))


def spline_w(table: np.ndarray, h: float, r: np.ndarray) -> np.ndarray:
    """Linear interpolation of the W table at distance r (sph_funcs.py:92-104)."""
    n = table.shape[0]
    scale = (n - 1) / h
    idx = r * scale
    lo = np.minimum(idx.astype(np.int64), n - 2)
    a = idx - lo
    w = (1 - a) * table[lo].astype(np.float64) + a * table[lo + 1].astype(np.float64)
    return np.where(r >= h, 0.0, w)


def features(cfg, dt, x, rho_adv, a_ii, ff_count, carry, nbr_idx, w_table) -> np.ndarray:
    """[n, 8] features; nbr_idx[i, :ff_count[i]] are i's fluid neighbours."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    f = np.zeros((n, 8))
    f[:, 0] = (np.asarray(rho_adv, np.float64) / cfg.rho0 - 1.0) * 100.0
    f[:, 1] = np.asarray(a_ii, np.float64) / a_ref(cfg, dt)
    f[:, 2] = np.asarray(ff_count, np.float64) / 32.0
    f[:, 3] = np.asarray(carry, np.float64) / p_ref(cfg)
    f[:, :4] = np.clip(f[:, :4], -FEAT_CLAMP, FEAT_CLAMP)
    w0 = spline_w(w_table, cfg.h, np.zeros(1))[0]
    for i in range(n):
        j = nbr_idx[i, :ff_count[i]]
        w = spline_w(w_table, cfg.h, np.linalg.norm(x[j] - x[i], axis=1))
        sw = w0 + w.sum()
        f[i, 4:] = (w0 * f[i, :4] + w @ f[j, :4]) / sw
    return f


def start_value(carry, y, a_ii, cfg, f1=None) -> np.ndarray:
    """max(0, carry + P_ref * y) in float32; the carried value where
    |f1| = |a_ii / a_ref| < 0.05; 0 where |a_ii| < inv_small."""
    carry = np.asarray(carry, np.float32)
    pn = carry + np.float32(p_ref(cfg)) * np.asarray(y, np.float32)
    pn = np.where(np.isfinite(pn), pn, carry)
    pn = np.maximum(pn, np.float32(0)).astype(np.float32)
    if f1 is not None:
        pn = np.where(np.abs(f1) < WEAK_COUPLING, carry, pn).astype(np.float32)
    a = np.asarray(a_ii, np.float32)
    return np.where((a > -cfg.inv_small) & (a < cfg.inv_small), np.float32(0), pn).astype(np.float32)
