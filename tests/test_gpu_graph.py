"""The step replayed from captured hipGraphs (graph_iters > 0) against the
same step issued kernel by kernel: identical states after every step (the
graphs are replayed many times, with host reads and synchronisations
between the replays) and identical Jacobi iteration counts, for the
previous-step start, the learned start, and a batch of scenes."""
import sys

import numpy as np
import pytest

from helpers import assert_same
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _pair(cfg, k_graph, net=None):
    from warpsph_amd.solver import IISPH
    sc = scene.build(cfg)
    a = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True, predictor=net)
    b = IISPH(stage_path=None, config=cfg.replace(graph_iters=k_graph), scene=sc,
              collect_stats=True, predictor=net)
    return a, b


@pytest.mark.parametrize("k_graph", [1, 2, 6])
def test_graph_step_equals_eager(k_graph):
    cfg = config.dambreak_8k().replace(jitter=0.01, warm_start=config.WARM_PREVIOUS)
    _announce(f"graph step, graph_iters={k_graph}")
    a, b = _pair(cfg, k_graph)
    for k in range(30):
        a.step()
        b.step()
        assert b.step_stats()["iterations"] == a.step_stats()["iterations"], k
        if k % 5 == 4:
            for name in ("x", "v", "p"):
                assert_same(f"{name} step {k}", getattr(b, name).cpu().numpy(),
                            getattr(a, name).cpu().numpy())
    assert b.counters() == a.counters()


def test_graph_step_zero_start_and_timing():
    from warpsph_amd import _native as N
    cfg = config.dambreak_8k().replace(warm_start=config.WARM_ZERO)
    _announce("graph step, zero start, timing on")
    a, b = _pair(cfg, 4)
    N.check("wsph_set_timing", b._lib.wsph_set_timing(b._h, 1))
    for _ in range(10):
        a.step()
        b.step()
    assert_same("x", b.x.cpu().numpy(), a.x.cpu().numpy())
    import ctypes as C
    tm = (C.c_double * 8)()
    N.check("wsph_last_timings", b._lib.wsph_last_timings(b._h, tm, 8))
    assert tm[6] == 10 and tm[4] > 0


def test_kernel_timing_mode():
    """Timing mode 2 records only the two events around the first
    iteration's J1 + J2 pair, on every 4th step, and leaves the phase and the
    per-kernel sums at zero; the device-driven loop
    paces itself on the pinned progress word, with the same results as an
    untimed solver."""
    from warpsph_amd import _native as N
    import ctypes as C
    cfg = config.dambreak_8k()
    _announce("timing mode 2 (kernel events only)")
    a, b = _pair(cfg, 0)
    with pytest.raises(RuntimeError, match="timing mode"):
        N.check("wsph_set_timing", b._lib.wsph_set_timing(b._h, 3))
    N.check("wsph_set_timing", b._lib.wsph_set_timing(b._h, 2))
    for _ in range(12):
        a.step()
        b.step()
    assert_same("x", b.x.cpu().numpy(), a.x.cpu().numpy())
    assert b.counters() == a.counters()
    tm = (C.c_double * 12)()
    N.check("wsph_last_timings", b._lib.wsph_last_timings(b._h, tm, 12))
    assert tm[6] == 12 and tm[4] == 0 and tm[0] == 0
    assert tm[9] == 0 and tm[11] == 3 and tm[10] > 0


def test_graph_step_learned_start():
    from warpsph_amd.mlp import PressureNet
    cfg = config.dambreak_8k().replace(jitter=0.01, warm_start=config.WARM_LEARNED)
    _announce("graph step, learned start")
    net = PressureNet.load(str(__import__("pathlib").Path(__file__).resolve().parents[1]
                               / "warpsph_amd" / "pressure_net.npy"), max_batch=1024)
    a, b = _pair(cfg, 3, net)
    for _ in range(20):
        a.step()
        b.step()
    assert_same("x", b.x.cpu().numpy(), a.x.cpu().numpy())
    assert [r["start_source"] for r in b.history()] == [r["start_source"] for r in a.history()]


def test_graph_step_batch():
    from warpsph_amd.batch import BatchIISPH
    cfgs = [c.replace(warm_start=config.WARM_PREVIOUS)
            for c in scene.batch_configs(config.batch_256())[:5]]
    _announce("graph step, batch of 5 scenes")
    scs = [scene.build(c) for c in cfgs]
    a = BatchIISPH(cfgs, scs)
    b = BatchIISPH([c.replace(graph_iters=2) for c in cfgs], scs)
    for _ in range(12):
        a.step()
        b.step()
        np.testing.assert_array_equal(b.scene_stats()["iterations"], a.scene_stats()["iterations"])
    assert_same("x", b.x.cpu().numpy(), a.x.cpu().numpy())
    assert_same("p", b.p.cpu().numpy(), a.p.cpu().numpy())
