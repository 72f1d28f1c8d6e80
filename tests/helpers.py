"""Shared helpers of the parity tests: build a GPU solver and the CPU oracle
on the same scene and compare their fields."""
from __future__ import annotations

import math

import numpy as np

from warpsph_amd import config, scene, tables
from warpsph_amd._native import make_params


def brute_force_lists(xq, xs, h, exclude_self):
    """O(nq*ns) neighbour sets with the identical float32 expression
    sqrt((dx*dx + dy*dy) + dz*dz) < h, dx = q - s (IISPH.py:71-72)."""
    xq = np.asarray(xq, np.float32)
    xs = np.asarray(xs, np.float32)
    hf = np.float32(h)
    out = []
    for i in range(xq.shape[0]):
        d = xq[i][None, :] - xs
        r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        m = r < hf
        if exclude_self:
            m[i] = False
        out.append(set(np.nonzero(m)[0].tolist()))
    return out


def oracle_setup(cfg, sc=None):
    sc = sc if sc is not None else scene.build(cfg)
    gd, bd = scene.grid_dims_for(cfg, sc)
    P = make_params(cfg, gd, bd)
    wt, gt = tables.make_tables(cfg.h, cfg.w_size, cfg.grad_size)
    return sc, P, wt, gt


def lattice_block_scene(cfg, cells=(10, 10, 10)):
    return scene.dam_break(cfg.replace(fluid_cells=cells))


def assert_same(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{name}: shape {a.shape} vs {b.shape}"
    if a.dtype.kind == "f":
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if bad.any():
            k = np.argwhere(bad)[0]
            diff = np.abs(a.astype(np.float64) - b.astype(np.float64))
            raise AssertionError(
                f"{name}: {int(bad.sum())} of {a.size} values differ, first at {tuple(k)}: "
                f"{a[tuple(k)]!r} vs {b[tuple(k)]!r}; max abs diff {np.nanmax(diff):.3e}")
    else:
        assert np.array_equal(a, b), f"{name}: integer arrays differ"


# ---- the cubic spline in closed form (float64), independent of the tables --
def kernel_sigma(h):
    """sigma = 8 / (pi h^3) (wp_sph/main.py:60)."""
    return 8.0 / (math.pi * h ** 3)


def kernel_w(r, h):
    """W(r) = sigma [6 (q^3 - q^2) + 1] for q <= 1/2, 2 sigma (1 - q)^3 for
    q < 1, else 0, q = r / h."""
    q = r / h
    s = kernel_sigma(h)
    if q <= 0.5:
        return s * (6 * (q ** 3 - q ** 2) + 1)
    if q < 1:
        return 2 * s * (1 - q) ** 3
    return 0.0


def kernel_dw(r, h):
    """|dW/dr| of the cubic spline."""
    q = r / h
    s = kernel_sigma(h) / h
    if q <= 0.5:
        return abs(s * (18 * q * q - 12 * q))
    if q < 1:
        return abs(-6 * s * (1 - q) ** 2)
    return 0.0


def kernel_grad(xi, xj, h):
    """grad_spline_W(x_j - x_i) (sph_funcs.py:105-118) in closed form: the
    table holds -W'(r) / r, so the value is |W'(r)| times the unit vector
    from x_i to x_j."""
    r = np.asarray(xj, np.float64) - np.asarray(xi, np.float64)
    n = float(np.linalg.norm(r))
    if not 1e-10 < n < h:
        return np.zeros(3)
    return kernel_dw(n, h) * r / n
