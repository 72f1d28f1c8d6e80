"""Poisson-disk surface sampler (wsph_poisson_surface, host code) and the
torus-and-plane scene: minimum distance, points on the surface, coverage
(no hole wider than the algorithm allows), determinism by seed."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from warpsph_amd import config, scene

R, r, C = 0.3, 0.1, (0.6, 0.105, 0.6)


def torus_dist(p):
    p = np.asarray(p, np.float64)
    rho = np.hypot(p[:, 0] - C[0], p[:, 2] - C[2])
    return np.abs(np.hypot(rho - R, p[:, 1] - C[1]) - r)


@pytest.mark.parametrize("shape", ["plane", "torus"])
def test_min_distance_and_surface(shape):
    d = 0.01
    if shape == "plane":
        pts = scene.poisson_surface("plane", (0, 0, 1, 0.5, 0.25), d, 3)
        on = np.abs(pts[:, 1] - 0.25) < 1e-7
        inside = (pts[:, 0] >= 0) & (pts[:, 0] <= 1) & (pts[:, 2] >= 0) & (pts[:, 2] <= 0.5)
        assert on.all() and inside.all()
        area = 0.5
    else:
        pts = scene.poisson_surface("torus", (*C, R, r), d, 3)
        assert torus_dist(pts).max() < 1e-6
        area = 4 * np.pi ** 2 * R * r
    nn, _ = cKDTree(pts.astype(np.float64)).query(pts.astype(np.float64), k=2)
    # float32 output of points placed at >= d in float64
    assert nn[:, 1].min() >= d * (1 - 1e-6)
    # Bridson's sets are close to maximal: density between random close
    # packing of disks and hexagonal packing
    density = pts.shape[0] / area * d * d
    assert 0.6 < density < 1.1547, density


def test_coverage_has_no_large_holes():
    d = 0.01
    pts = scene.poisson_surface("plane", (0, 0, 1, 1, 0), d, 5)
    rng = np.random.default_rng(0)
    probe = np.stack([rng.uniform(0.02, 0.98, 20000), np.zeros(20000),
                      rng.uniform(0.02, 0.98, 20000)], 1)
    dist, _ = cKDTree(pts).query(probe)
    # every candidate annulus [d, 2d) around each sample was tried 30 times
    assert np.quantile(dist, 0.999) < 2 * d
    assert dist.max() < 3 * d


def test_seed_determinism():
    a = scene.poisson_surface("torus", (*C, R, r), 0.02, 7)
    b = scene.poisson_surface("torus", (*C, R, r), 0.02, 7)
    c = scene.poisson_surface("torus", (*C, R, r), 0.02, 8)
    np.testing.assert_array_equal(a, b)
    assert a.shape != c.shape or not np.array_equal(a, c)


def test_bad_arguments():
    import ctypes as Cc
    from warpsph_amd import _native as N
    lib = N.load()
    cnt = Cc.c_int()
    prm = (Cc.c_double * 5)(0, 0, 1, 1, 0)
    assert lib.wsph_poisson_surface(5, prm, 0.01, 1, 0, None, Cc.byref(cnt)) == -1
    assert lib.wsph_poisson_surface(0, prm, -1.0, 1, 0, None, Cc.byref(cnt)) == -1
    out = np.zeros((3, 3), np.float32)
    assert lib.wsph_poisson_surface(0, prm, 0.1, 1, 3, out.ctypes.data, Cc.byref(cnt)) == -3


def test_torus_workload_shape():
    cfg = config.torus_4m()
    sc = scene.build(cfg)
    assert sc.n == 160 ** 3
    # plane first (hidden in renders), then the torus
    assert 90000 < sc.boundary_hide_n < 110000
    assert 25000 < sc.nb - sc.boundary_hide_n < 35000
    assert torus_dist(sc.xb[sc.boundary_hide_n:]).max() < 1e-6
    # the fluid starts above the torus
    assert sc.x[:, 1].min() > C[1] + r
    gd, _ = scene.grid_dims_for(cfg, sc)
    assert all(3 <= v <= 1024 for v in gd)
