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


# ---- hydrostatic pressure of a settled layer -----------------------------------
def hydrostatic_config():
    """A fluid layer filling the floor of a 0.4 x 0.6 x 0.4 m box: 20 x 15 x 20
    particles at 0.02 m (a 0.30 m column), one wall layer, g = 9.81, dt = 1e-3,
    XSPH 0.01, previous-step start. It settles without a dam-break collapse."""
    return config.dambreak_8k().replace(box=(0.4, 0.6, 0.4), fluid_cells=(20, 15, 20),
                                        warm_start=config.WARM_PREVIOUS)


def hydrostatic_fit(cfg, x, p):
    """Least-squares p = a y + b over the particles at least 2 spacings from
    the side walls and the floor and 3 below the top; returns (a, -b / a), the
    slope (hydrostatics: -rho0 g) and the height where the fit crosses p = 0."""
    x = np.asarray(x, np.float64)
    p = np.asarray(p, np.float64)
    d = cfg.spacing
    bx, _, bz = cfg.box
    inner = ((x[:, 0] > 2 * d) & (x[:, 0] < bx - 3 * d) & (x[:, 2] > 2 * d)
             & (x[:, 2] < bz - 3 * d))
    top = np.percentile(x[:, 1], 99.5)
    deep = inner & (x[:, 1] > 2 * d) & (x[:, 1] < top - 3 * d)
    A = np.vstack([x[deep, 1], np.ones(int(deep.sum()))]).T
    a, b = np.linalg.lstsq(A, p[deep], rcond=None)[0]
    return a, -b / a


def assert_hydrostatic(cfg, slopes, surfaces):
    """The time-averaged pressure gradient of the settled layer is rho0 g
    within 5%, and its zero crossing lies within one spacing of the column
    height (15 layers of 0.02 m filling from y = -0.01, the wall's half
    spacing: 0.29 m)."""
    rho_g = cfg.rho0 * cfg.gravity
    slope = float(np.mean(slopes))
    surface = float(np.mean(surfaces))
    assert abs(-slope - rho_g) < 0.05 * rho_g, (slope, rho_g)
    assert abs(surface - 0.29) < cfg.spacing, surface
