"""wsph_mesh_inside (the containment test of model.py:14-47) on the GPU,
bit for bit against the oracle, and a mesh-sampled model moving through the
fluid against the oracle solver."""
import sys

import numpy as np
import pytest

from helpers import assert_same
from warpsph_amd import config, mesh as M, scene

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _axes(lo, hi, d):
    return [np.arange(lo[a], hi[a], d) for a in range(3)]


def _noisy_sphere():
    v, t = M.icosphere((0.3, 0.2, 0.25), 0.2, subdiv=3)
    rng = np.random.default_rng(5)
    return v + rng.uniform(-0.01, 0.01, v.shape), t


CASES = {
    "box aligned": (M.box_mesh((0.1, 0.2, 0.3), (0.7, 0.6, 0.9)), _axes((0, 0, 0), (1, 1, 1), 0.05)),
    "box unaligned": (M.box_mesh((0.1, 0.2, 0.3), (0.73, 0.61, 0.95)),
                      _axes((-0.013, 0.011, 0.017), (0.9, 0.8, 1.1), 0.05)),
    "octahedron": (M.octahedron((0, 0, 0), 1.0), [np.arange(-1.5, 1.5001, 0.125)] * 3),
    "icosphere": (M.icosphere((0.5, 0.4, 0.45), 0.3, 2), _axes((0.1, 0, 0.05), (0.9, 0.8, 0.85), 0.021)),
    "torus": (M.torus_mesh((0, 0, 0), 0.3, 0.1, 40, 20), _axes((-0.45, -0.15, -0.45), (0.45, 0.15, 0.45), 0.013)),
    "noisy sphere": (_noisy_sphere(), _axes((0.05, -0.05, 0.0), (0.55, 0.45, 0.5), 0.009)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_inside_matches_oracle(orc, name):
    (v, t), axes = CASES[name]
    _announce(f"mesh: {name}, {t.shape[0]} triangles, {np.prod([a.size for a in axes])} points")
    got = M.inside_mask((v, t), *axes)
    want = orc.mesh_inside(v, t, *axes)
    assert want.any()
    assert np.array_equal(got, want), f"{int((got != want).sum())} points differ"


def test_large_lattice_against_oracle_subset(orc):
    """A 2.2M-point lattice in a 4,608-triangle torus: each point's result
    depends on that point only, so the oracle checks every 9th column in x
    and z (all heights)."""
    v, t = M.torus_mesh((0, 0, 0), 0.3, 0.1, 96, 24)
    axes = _axes((-0.42, -0.12, -0.42), (0.42, 0.12, 0.42), 0.0051)
    _announce(f"mesh: large torus lattice {[a.size for a in axes]}")
    got = M.inside_mask((v, t), *axes)
    sub = [axes[0][::9], axes[1], axes[2][::9]]
    want = orc.mesh_inside(v, t, *sub)
    assert np.array_equal(got[::9, :, ::9], want)


def test_sample_points_and_load_model(orc):
    _announce("mesh: load_model pipeline")
    spec = {"mesh": "torus", "center": (0.0, 0.0, 0.0), "R": 0.1, "r": 0.04}
    rot = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0]])
    pos = np.array([0.6, 0.16, 0.5])
    got = M.load_model(spec, 1.5, pos, rot, 0.02)
    (v, t), axes = M.lattice_axes(spec, 1.5, 0.02)
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    want = grid[orc.mesh_inside(v, t, *axes).reshape(-1)] @ rot + pos
    assert got.shape[0] > 100
    assert np.array_equal(got, want)


def test_bad_index_and_empty_inputs():
    _announce("mesh: argument checks")
    from warpsph_amd._native import WsphError
    with pytest.raises(WsphError, match="index"):
        M.inside_mask((np.zeros((3, 3)), np.array([[0, 1, 3]], np.int32)), [0.0], [0.0], [0.0])
    empty = (np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert not M.inside_mask(empty, [0.0, 1.0], [0.0], [0.0, 1.0]).any()
    v, t = M.box_mesh((0, 0, 0), (1, 1, 1))
    assert M.inside_mask((v, t), [], [0.5], [0.5]).shape == (0, 1, 1)


def test_mesh_model_scene_matches_oracle(orc):
    """wheel_8k: a mesh-sampled torus moving into the dam break; the solver
    follows the oracle bit for bit, boundary positions included."""
    _announce("mesh: wheel_8k scene, 30 steps")
    from warpsph_amd.solver import IISPH
    cfg = config.wheel_8k().replace(warm_start=config.WARM_PREVIOUS, rigid_v=(-2.0, 0.0, 0.0))
    sc = scene.build(cfg)
    assert sc.nb - sc.boundary_hide_n > 100
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    o = orc.OracleSolver(cfg, sc)
    for _ in range(30):
        s.step()
        info = o.step()
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("p", s.p.cpu().numpy(), info["p"])
    assert_same("boundary x", s.boundary_x.cpu().numpy(), o.xb)
