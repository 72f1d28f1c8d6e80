"""Batched scenes (wsph_set_scenes): every scene of a batch evolves exactly as
it does in a solver of its own — positions, velocities and pressures equal
bit for bit, and each scene's Jacobi iteration count per step equals its solo
count (its own stop test) — with the device-driven and the persistent Jacobi
loop."""
import sys

import numpy as np
import pytest

from helpers import assert_same
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _scene_cfgs(k=4, max_n=9000):
    out = []
    for c in scene.batch_configs(config.batch_256()):
        n = c.fluid_cells[0] * c.fluid_cells[1] * c.fluid_cells[2]
        if n <= max_n:
            out.append(c)
        if len(out) == k:
            return out
    raise AssertionError("not enough small scenes")


@pytest.mark.parametrize("jacobi,start", [
    (config.JACOBI_DEVICE, config.WARM_PREVIOUS),
    (config.JACOBI_PERSISTENT, config.WARM_PREVIOUS),
    (config.JACOBI_DEVICE, config.WARM_ZERO),
])
def test_batch_equals_solo_runs(jacobi, start):
    from warpsph_amd.batch import BatchIISPH
    from warpsph_amd.solver import IISPH
    steps = 12
    cfgs = [c.replace(jacobi_mode=jacobi, warm_start=start) for c in _scene_cfgs()]
    _announce(f"batch of {len(cfgs)} scenes, jacobi={jacobi} start={start}")
    scs = [scene.build(c) for c in cfgs]
    b = BatchIISPH(cfgs, scs)
    per_step = []
    for _ in range(steps):
        b.step()
        per_step.append(b.scene_stats()["iterations"].copy())
    per_step = np.array(per_step)
    xb, vb, pb = (b.split(t.cpu().numpy()) for t in (b.x, b.v, b.p))
    for k, (c, sc) in enumerate(zip(cfgs, scs)):
        s = IISPH(stage_path=None, config=c, scene=sc, collect_stats=True)
        iters = []
        for _ in range(steps):
            s.step()
            iters.append(s.step_stats()["iterations"])
        assert_same(f"scene {k} x", xb[k], s.x.cpu().numpy())
        assert_same(f"scene {k} v", vb[k], s.v.cpu().numpy())
        assert_same(f"scene {k} p", pb[k], s.p.cpu().numpy())
        assert list(per_step[:, k]) == iters, (k, list(per_step[:, k]), iters)
        s.close()
    # the batch stops when its slowest scene does
    c = b.counters()
    assert c["iterations"] == int(per_step.max(axis=1).sum())
    b.close()


def test_batch_of_one_scene_matches_solo():
    from warpsph_amd.batch import BatchIISPH
    from warpsph_amd.solver import IISPH
    cfg = config.dambreak_8k().replace(jitter=0.01, warm_start=config.WARM_PREVIOUS)
    _announce("batch of one scene")
    sc = scene.build(cfg)
    b = BatchIISPH([cfg], [sc])
    s = IISPH(stage_path=None, config=cfg, scene=sc)
    for _ in range(8):
        b.step()
        s.step()
    assert_same("x", b.x.cpu().numpy(), s.x.cpu().numpy())
    b.close()
    s.close()


@pytest.mark.parametrize("jacobi,start", [
    (config.JACOBI_DEVICE, config.WARM_PREVIOUS),
    (config.JACOBI_PERSISTENT, config.WARM_PREVIOUS),
    (config.JACOBI_DEVICE, config.WARM_ZERO),
])
def test_batch_scenes_follow_oracle(orc, jacobi, start):
    """Each scene of a 6-scene batch drawn from the C3 scene generator
    (batch_configs) follows the CPU oracle's IISPH.step (IISPH.py:1002-1371)
    bit for bit: x, v, p after every step and the scene's Jacobi iterations
    per step."""
    from warpsph_amd.batch import BatchIISPH
    steps = 10
    cfgs = [c.replace(jacobi_mode=jacobi, warm_start=start) for c in _scene_cfgs(6, 12000)]
    _announce(f"batch of {len(cfgs)} scenes against the oracle, jacobi={jacobi} start={start}")
    scs = [scene.build(c) for c in cfgs]
    b = BatchIISPH(cfgs, scs)
    orcs = [orc.OracleSolver(c, sc) for c, sc in zip(cfgs, scs)]
    for k in range(steps):
        b.step()
        iters = b.scene_stats()["iterations"]
        xs, vs, ps = (b.split(t.cpu().numpy()) for t in (b.x, b.v, b.p))
        for q, o in enumerate(orcs):
            info = o.step()
            assert iters[q] == info["iterations"], (k, q, iters[q], info["iterations"])
            assert_same(f"step {k} scene {q} x", xs[q], o.x)
            assert_same(f"step {k} scene {q} v", vs[q], o.v)
            assert_same(f"step {k} scene {q} p", ps[q], o.p)
    b.close()
