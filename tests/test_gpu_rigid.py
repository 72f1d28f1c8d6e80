"""Moving rigid boundary (the reference's PLANE / HAND scenes: set_rigid_v at
sim_time 0, update_rigid_x after the drift, the boundary grid rebuilt every
step, IISPH.py:652-671, 1061-1081, 1355-1364, 1415-1416) on the HIP path,
compared bit for bit with the oracle: fluid state, pressure and the moved
boundary positions, under each Jacobi driver."""
import sys

import numpy as np
import pytest

from helpers import assert_same
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _solver(cfg, sc):
    from warpsph_amd.solver import IISPH
    return IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)


def _compare(s, o, info, what):
    assert_same(f"{what} x", s.x.cpu().numpy(), o.x)
    assert_same(f"{what} v", s.v.cpu().numpy(), o.v)
    assert_same(f"{what} p", s.p.cpu().numpy(), info["p"])
    assert_same(f"{what} boundary x", s.boundary_x.cpu().numpy(), o.xb)
    assert_same(f"{what} boundary v", s.boundary_v.cpu().numpy(), o.vb)


DRIVERS = {
    "device": dict(jacobi_mode=config.JACOBI_DEVICE),
    "host": dict(jacobi_mode=config.JACOBI_HOST),
    "persistent": dict(jacobi_mode=config.JACOBI_PERSISTENT),
    "graph": dict(jacobi_mode=config.JACOBI_DEVICE, graph_iters=2, chunk_iters=2),
}


@pytest.mark.parametrize("driver", list(DRIVERS))
def test_paddle_matches_oracle(orc, driver):
    steps = 40 if driver == "device" else 12
    _announce(f"rigid: paddle_8k, {driver} driver, {steps} steps")
    cfg = config.paddle_8k().replace(warm_start=config.WARM_PREVIOUS, **DRIVERS[driver])
    sc = scene.build(cfg)
    s = _solver(cfg, sc)
    o = orc.OracleSolver(cfg, sc)
    for k in range(steps):
        s.step()
        info = o.step()
        if k in (0, 1, steps // 2):
            _compare(s, o, info, f"step {k}")
    _compare(s, o, info, f"step {steps - 1}")
    # the paddle has moved into the fluid and the fluid feels it
    hide = sc.boundary_hide_n
    moved = s.boundary_x.cpu().numpy()[hide:] - sc.xb[hide:]
    assert np.all(moved[:, 0] < -0.01) and np.all(moved[:, 1:] == 0)
    assert s.fs_neighbor_num.cpu().numpy().max() > 0


def test_static_rigid_equals_static_boundary(orc):
    """first = nb: nothing moves, but the boundary is re-sorted every step;
    the result equals the static solver's, bit for bit. A rigid motion set
    after the first step never receives its velocity (the reference sets it
    only at sim_time 0), so it moves nothing either."""
    _announce("rigid: no-op rigid motion equals the static boundary")
    cfg = config.paddle_8k().replace(rigid_v=None, warm_start=config.WARM_PREVIOUS)
    sc = scene.build(cfg)
    a = _solver(cfg, sc)
    b = _solver(cfg, sc)
    c = _solver(cfg, sc)
    b.set_rigid(sc.nb, (5.0, 0.0, 0.0))
    for k in range(8):
        a.step()
        b.step()
        c.step()
        if k == 0:
            c.set_rigid(sc.boundary_hide_n, (-1.0, 0.0, 0.0))
    for other in (b, c):
        assert_same("x", other.x.cpu().numpy(), a.x.cpu().numpy())
        assert_same("p", other.p.cpu().numpy(), a.p.cpu().numpy())
        assert_same("boundary x", other.boundary_x.cpu().numpy(), sc.xb)
        assert_same("psi", other.boundary_phi.cpu().numpy(), a.boundary_phi.cpu().numpy())


def test_rigid_clears_with_new_boundary():
    _announce("rigid: wsph_set_rigid argument checks")
    from warpsph_amd import _native as N
    cfg = config.paddle_8k()
    s = _solver(cfg, scene.build(cfg))
    with pytest.raises(RuntimeError, match="rigid velocity"):
        N.check("wsph_set_rigid", s._lib.wsph_set_rigid(s._h, 0, None))
    s.set_rigid(-1, None)  # static again
    s.step()
    assert_same("boundary x", s.boundary_x.cpu().numpy(), s.scene.xb)


def test_reference_plane_scene_matches_oracle(orc, tmp_path, monkeypatch):
    """The reference's PLANE scene through wp_sph.solver.IISPH with
    wp_sph.main.scene_type switched, at a coarse spacing, its model a box mesh
    standing in for the missing plane.obj: the model moves at (2, 0, 0) and
    the run follows the oracle bit for bit."""
    _announce("rigid: reference PLANE scene with a stand-in model")
    import wp_sph.main as wm
    from wp_sph.solver import IISPH as RefIISPH
    from warpsph_amd import mesh as M
    v, t = M.box_mesh((0.0, 0.0, 0.0), (400.0, 100.0, 600.0))
    lines = [f"v {a} {b} {c}" for a, b, c in v] + [f"f {a + 1} {b + 1} {c + 1}" for a, b, c in t]
    (tmp_path / "plane.obj").write_text("\n".join(lines) + "\n")
    monkeypatch.setattr(wm, "scene_type", wm.SceneType.PLANE)
    monkeypatch.setattr(wm, "MODEL_DIR", tmp_path)
    cfg = config.reference("PLANE", models_dir=str(tmp_path)).replace(spacing=0.25)
    sc = scene.build(cfg)
    assert sc.boundary_hide_n == 0 and sc.nb > 20
    s = RefIISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    o = orc.OracleSolver(cfg, sc)
    for _ in range(6):
        s.step()
        info = o.step()
    _compare(s, o, info, "PLANE")
    assert np.all(s.boundary_x.cpu().numpy()[:, 0] > sc.xb[:, 0])
    # without a config the class builds the scene scene_type names
    s2 = RefIISPH(stage_path=None)
    assert s2.cfg.rigid_v == (2.0, 0.0, 0.0) and s2.boundary_hide_n == 0
