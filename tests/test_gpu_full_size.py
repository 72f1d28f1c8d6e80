"""Bit-exact parity at full size against the CPU oracle: the
1,000,000-particle dam break (three steps), the 15,625,000-particle one at
0.004 m (one step), the 1M scene with a moving paddle (two steps), the
4,096,000-particle torus-and-plane scene (one step), all 256 scenes of the
batched workload (two steps), and the 1M dam break with the learned start run
by a constant network (three steps). The Jacobi
iteration count is forced (4 per step) so that the oracle, OpenMP on the
host, finishes in seconds; the stop rule is compared at the smaller sizes
(tests/test_gpu_parity.py). This covers what only full size exercises: list
capacities near their use, 64-bit list offsets past 4 GiB (16M particles x 128
entries x 4 B), the hash grid at 260 x 135 x 135 cells, and block counts in
the tens of thousands."""
import sys

import numpy as np
import pytest

from helpers import assert_same
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("preset,steps", [("dambreak_1m", 3), ("dambreak_16m", 1), ("paddle_1m", 2)])
def test_full_size_matches_oracle(orc, preset, steps):
    from warpsph_amd.solver import IISPH
    cfg = config.preset(preset).replace(warm_start=config.WARM_PREVIOUS, force_iters=4,
                                        block=512, cell_layout=config.LAYOUT_MORTON)
    sc = scene.build(cfg)
    _announce(f"full size {preset}: {sc.n} fluid, {sc.nb} boundary")
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    o = orc.OracleSolver(cfg, sc)
    info = None
    for k in range(steps):
        _announce(f"full size {preset}: step {k + 1}")
        s.step()
        info = o.step()
        assert s.step_stats()["iterations"] == info["iterations"] == 4
    fields = dict(rho=s.rho, rho_adv=s.rho_adv, a_ii=s.term_a_ii, p=s.p, Ap=s.term_Ap_i,
                  ff_count=s.ff_neighbor_num, fs_count=s.fs_neighbor_num)
    for name, t in fields.items():
        assert_same(name, t.cpu().numpy(), info[name])
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("v", s.v.cpu().numpy(), o.v)
    assert np.all(np.isfinite(o.x))
    s.close()


def test_torus_4m_matches_oracle(orc):
    """The 4,096,000-particle torus-and-plane scene (no walls, Poisson-disk
    sampled surfaces, boundary grid apart from the fluid's): one step."""
    from warpsph_amd.solver import IISPH
    cfg = config.torus_4m().replace(warm_start=config.WARM_PREVIOUS, force_iters=4)
    sc = scene.build(cfg)
    _announce(f"full size torus_4m: {sc.n} fluid, {sc.nb} boundary")
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    o = orc.OracleSolver(cfg, sc)
    s.step()
    info = o.step()
    for name, t in dict(rho=s.rho, p=s.p, ff_count=s.ff_neighbor_num,
                        fs_count=s.fs_neighbor_num).items():
        assert_same(name, t.cpu().numpy(), info[name])
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("v", s.v.cpu().numpy(), o.v)
    s.close()


def test_c3_batch_matches_oracle(orc):
    """All 256 scenes of the batched workload (4,747,246 fluid particles in one
    batched solver), two steps, each scene against its own oracle run."""
    from warpsph_amd.batch import BatchIISPH
    cfgs = [c.replace(warm_start=config.WARM_PREVIOUS, force_iters=4)
            for c in scene.batch_configs(config.batch_256())]
    scs = [scene.build(c) for c in cfgs]
    _announce(f"full size batch: {len(cfgs)} scenes, {sum(sc.n for sc in scs)} fluid")
    b = BatchIISPH(cfgs, scs)
    orcs = [orc.OracleSolver(c, sc) for c, sc in zip(cfgs, scs)]
    for k in range(2):
        _announce(f"full size batch: step {k + 1}")
        b.step()
        xs, vs, ps = (b.split(t.cpu().numpy()) for t in (b.x, b.v, b.p))
        for q, o in enumerate(orcs):
            o.step()
            assert_same(f"step {k} scene {q} x", xs[q], o.x)
            assert_same(f"step {k} scene {q} v", vs[q], o.v)
            assert_same(f"step {k} scene {q} p", ps[q], o.p)
    b.close()


def test_learned_start_constant_network_1m(orc):
    """The learned start at 1M with a constant network (every weight 0, output
    bias b3): the matrix cores return the bias exactly, so the start value is
    max(0, carry + P_ref b3), and the oracle, handed that start (or the carried
    pressure, as the arbiter chose), reproduces the trajectory bit for bit.
    This runs the feature gather, the inline MFMA inference, the start's
    guards and the arbiter at full size."""
    import learn_oracle as LO
    import mlp_oracle as MO
    from warpsph_amd.mlp import PressureNet
    from warpsph_amd.solver import IISPH
    b3, steps = 0.05, 3
    p = np.zeros(MO.N_PARAM, np.float32)
    p[MO.OFF_B3] = b3
    net = PressureNet(max_batch=1024, params=p)
    cfg = config.dambreak_1m().replace(warm_start=config.WARM_LEARNED, force_iters=4)
    sc = scene.build(cfg)
    _announce("full size learned start, constant network, 1M")
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True, predictor=net)
    for _ in range(steps):
        s.step()
    hist = s.history()[-steps:]
    o = orc.OracleSolver(cfg.replace(warm_start=config.WARM_PREVIOUS), sc)
    carry = np.zeros(sc.n, np.float32)
    for k in range(steps):
        _announce(f"full size learned start: oracle step {k + 1}")
        net_start = LO.start_value(carry, np.float32(b3), np.ones_like(carry), cfg)
        p0 = net_start if hist[k]["start_source"] == config.WARM_LEARNED else carry
        info = o.step(p0=p0)
        # no particle is weakly coupled or inside the a_ii guard band in these steps,
        # so the start does not depend on a_ii
        assert np.all(np.abs(info["a_ii"] / LO.a_ref(cfg, cfg.dt0)) >= LO.WEAK_COUPLING)
        assert hist[k]["iterations"] == info["iterations"] == 4
        carry = info["p_carry"]
    assert hist[0]["start_source"] == config.WARM_LEARNED
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("v", s.v.cpu().numpy(), o.v)
    assert_same("p", s.p.cpu().numpy(), info["p"])
    s.close()
