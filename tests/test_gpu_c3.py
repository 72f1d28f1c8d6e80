"""The batched workload at its full size: 256 independent dam-break scenes
(fluid counts log-uniform in 2k-64k, boxes 0.5-2 m, seed 2; 4,747,246 fluid
particles) solved in one batched solver, each scene converging to its own
tolerance (the stop test of IISPH.py:1200-1207 per scene).

* 20 steps with the previous-step start and 20 with the learned start: every
  scene converges every step, within the reference's 2..101 iterations;
* with the previous-step start, one sampled scene equals its own solo solver
  bit for bit (x, v, p and its iterations per step; the learned start's
  arbiter decides for a whole solver, so a batch and a solo run may start
  differently);
* the logged training pairs: per step, min(per_step, n) rows of every scene,
  all finite.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import assert_same
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu

STEPS = 20
SAMPLE = 37  # the scene compared with its solo run


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c3():
    cfgs = scene.batch_configs(config.batch_256())
    scs = [scene.build(c) for c in cfgs]
    return cfgs, scs


@pytest.mark.parametrize("start", [config.WARM_PREVIOUS, config.WARM_LEARNED])
def test_c3_every_scene_converges(c3, start):
    from warpsph_amd.batch import BatchIISPH
    from warpsph_amd.mlp import PressureNet
    from warpsph_amd.solver import IISPH
    cfgs, scs = c3
    cfgs = [c.replace(warm_start=start) for c in cfgs]
    net = PressureNet.load(str(Path(__file__).resolve().parents[1] / "warpsph_amd" /
                               "pressure_net.npy")) if start == config.WARM_LEARNED else None
    n = sum(s.n for s in scs)
    _announce(f"C3: {len(cfgs)} scenes, {n} fluid particles, start={start}")
    assert len(cfgs) == 256 and n == 4747246
    b = BatchIISPH(cfgs, scs, predictor=net)
    iters = []
    for k in range(STEPS):
        st = b.step(stats=True)
        ss = b.scene_stats()
        assert st["status"] == 0, st
        assert ss["converged"].all(), (k, np.nonzero(~ss["converged"])[0])
        assert (ss["residual"] <= cfgs[0].eta).all()
        assert ((ss["iterations"] >= 2) & (ss["iterations"] <= 101)).all(), ss["iterations"]
        iters.append(ss["iterations"].copy())
    iters = np.array(iters)
    c = b.counters()
    assert c["nonconverged"] == 0 and c["steps"] == STEPS
    assert c["iterations"] == int(iters.max(axis=1).sum())
    xs, vs, ps = (b.split(t.cpu().numpy()) for t in (b.x, b.v, b.p))
    _announce(f"C3 start={start}: mean scene iterations {iters.mean():.2f}, "
              f"batch iterations/step {c['iterations'] / STEPS:.2f}")
    b.close()
    if start != config.WARM_PREVIOUS:
        return  # the learned start's arbiter decides for the whole solver
    s = IISPH(stage_path=None, config=cfgs[SAMPLE], scene=scs[SAMPLE], collect_stats=True)
    solo = []
    for _ in range(STEPS):
        s.step()
        solo.append(s.step_stats()["iterations"])
    assert list(iters[:, SAMPLE]) == solo
    assert_same("x", xs[SAMPLE], s.x.cpu().numpy())
    assert_same("v", vs[SAMPLE], s.v.cpu().numpy())
    assert_same("p", ps[SAMPLE], s.p.cpu().numpy())
    s.close()


def test_c3_logged_pairs(c3):
    from warpsph_amd import learn
    cfgs, scs = c3
    per_step = 64
    _announce("C3 pair logging")
    X, T = learn.generate_pairs_batched(cfgs, STEPS, per_step, seed=3)
    rows = STEPS * sum(min(per_step, s.n) for s in scs)
    assert X.shape == (rows, 8) and T.shape == (rows,)
    assert bool(X.isfinite().all()) and bool(T.isfinite().all())
