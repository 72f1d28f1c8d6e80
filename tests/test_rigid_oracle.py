"""The oracle's moving rigid boundary (set_rigid_v / update_rigid_x /
DYNAMIC_SCENE, IISPH.py:652-671, 1061-1081, 1355-1364, 1415-1416), pinned by
a closed form and by brute force. CPU only."""
import numpy as np
import pytest

from helpers import brute_force_lists
from warpsph_amd import config, scene


def _small_paddle(**kw):
    cfg = config.paddle_8k().replace(fluid_cells=(8, 8, 8), **kw)
    return cfg, scene.build(cfg)


def test_rigid_drift_closed_form(orc):
    """dt = 2^-10 and v = (-1, 0.5, 0) make every float32 update exact, so
    after k steps the model sits at x0 + k v dt (exactly), the walls have not
    moved, and the velocity was set in the first step (it moves in step 0)."""
    dt = 2.0 ** -10
    v = (-1.0, 0.5, 0.0)
    cfg, sc = _small_paddle(dt=dt, rigid_v=v)
    o = orc.OracleSolver(cfg, sc)
    hide = sc.boundary_hide_n
    x0 = sc.xb.astype(np.float64)
    for k in range(1, 6):
        o.step()
        want = x0[hide:] + k * np.asarray(v) * dt
        assert np.array_equal(o.xb[hide:].astype(np.float64), want), f"step {k}"
        assert np.array_equal(o.xb[:hide], sc.xb[:hide])
    assert np.array_equal(o.vb[hide:], np.broadcast_to(np.float32(v), o.vb[hide:].shape))
    assert not o.vb[:hide].any()


def test_velocity_set_only_at_time_zero(orc):
    """A rigid motion attached after the first step never gets its velocity
    (the reference's set_rigid_v runs when sim_time == 0)."""
    cfg, sc = _small_paddle(rigid_v=None)
    o = orc.OracleSolver(cfg, sc)
    o.step()
    o.rigid_first = sc.boundary_hide_n
    o.rigid_v = np.float32([-1.0, 0.0, 0.0])
    for _ in range(3):
        o.step()
    assert np.array_equal(o.xb, sc.xb)


def test_boundary_neighbours_follow_moved_model(orc):
    """Each step's fluid-boundary neighbour counts equal brute force over the
    boundary positions at the start of that step, so the boundary grid is
    rebuilt from the moved positions. A fast paddle (4 m/s) starts beyond h
    of the fluid and is inside it after a few steps."""
    cfg, sc = _small_paddle(rigid_v=(-4.0, 0.0, 0.0),
                            extra={"model": {"kind": "lattice", "cells": (2, 6, 6),
                                             "origin": (0.2, 0.0, 0.04)}})
    o = orc.OracleSolver(cfg, sc)
    h = cfg.h
    touched = 0
    for k in range(12):
        xf, xb = o.x.copy(), o.xb.copy()
        info = o.step()
        want = np.array([len(s) for s in brute_force_lists(xf, xb, h, False)], np.int32)
        assert np.array_equal(info["fs_count"], want), f"step {k}"
        paddle = brute_force_lists(xf, xb[sc.boundary_hide_n:], h, False)
        touched = max(touched, sum(1 for s in paddle if s))
    assert touched > 0


def test_batch_rejects_rigid_motion():
    from warpsph_amd.batch import BatchIISPH
    cfg = config.paddle_8k()
    with pytest.raises(ValueError, match="rigid"):
        BatchIISPH([cfg, cfg])


def test_reference_scene_types():
    """The reference's scene types (main.py:13-53, IISPH.py:751-766, 951-985):
    EMPTY is the default scene; PLANE has no walls and no gravity and moves its
    model along +x; HAND's model is turned 160 degrees about z and moves up;
    HOUSE keeps a static model."""
    import numpy as np
    assert config.reference("EMPTY") == config.reference()
    assert scene.box_shell((8.0, 8.0, 8.0), 0.05, 0).shape == (0, 3)
    pl = config.reference("PLANE", models_dir="/models")
    assert pl.boundary_layers == 0 and pl.gravity == 0.0 and pl.rigid_v == (2.0, 0.0, 0.0)
    assert pl.extra["model"]["mesh"] == "/models/plane.obj"
    hand = config.reference("HAND")
    R = np.array(hand.extra["model"]["rot"])
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15) and abs(np.linalg.det(R) - 1) < 1e-15
    assert np.allclose(np.array([1.0, 0.0, 0.0]) @ R,
                       [np.cos(np.radians(160)), -np.sin(np.radians(160)), 0.0])
    assert hand.rigid_v == (0.0, 0.1, 0.0)
    assert config.reference("HOUSE").rigid_v is None
    with pytest.raises(ValueError):
        config.reference("CASTLE")
