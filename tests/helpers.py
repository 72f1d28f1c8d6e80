"""Shared helpers of the parity tests: build a GPU solver and the CPU oracle
on the same scene and compare their fields."""
from __future__ import annotations

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
