"""Launch-shape autotuner (warpsph_amd/autotune.py): candidate space, choice
of the fastest shape, per-device/workload cache. A fake solver whose step
time depends on the shape stands in for the GPU."""
import json
import time

import pytest

from warpsph_amd import autotune, config


class FakeSolver:
    def __init__(self, cfg, cost):
        self.cfg = cfg
        self.cost = cost

    def step(self):
        time.sleep(self.cost(self.cfg))

    def close(self):
        pass


@pytest.fixture
def no_gpu(monkeypatch):
    monkeypatch.setattr(autotune, "_device_name", lambda: "fake-device")
    import torch
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)


def test_candidate_space():
    c = autotune.candidates()
    assert len(c) == 3 * 2 * 7
    assert all(set(autotune.SHAPE_KEYS) == set(x) for x in c)
    drivers = {(x["jacobi_mode"], x["chunk_iters"], x["coop_blocks_per_cu"], x["graph_iters"]) for x in c}
    assert (config.JACOBI_PERSISTENT, 0, 2, 0) in drivers
    assert (config.JACOBI_DEVICE, 2, 0, 4) in drivers


def test_picks_fastest_and_caches(tmp_path, no_gpu):
    cands = [dict(block=b, cell_layout=1, jacobi_mode=2, chunk_iters=1, coop_blocks_per_cu=0,
                  graph_iters=0) for b in (128, 256, 512)]
    cost = lambda cfg: 0.004 if cfg.block == 256 else 0.008
    made = []

    def make(cfg):
        made.append(cfg.block)
        return FakeSolver(cfg, cost)

    cache = tmp_path / "tune.json"
    r = autotune.tune(make, config.dambreak_8k(), key="w", cands=cands, steps=3, warmup=1,
                      cache=cache, finalists=0)
    assert r["shape"]["block"] == 256 and not r["cached"]
    assert len(r["table"]) == 3 and made == [128, 256, 512]
    data = json.loads(cache.read_text())
    assert "fake-device|w" in data
    # a second call reads the cache and builds nothing
    made.clear()
    r2 = autotune.tune(make, config.dambreak_8k(), key="w", cands=cands, cache=cache)
    assert r2["cached"] and r2["shape"] == r["shape"] and made == []
    # force re-measures
    r3 = autotune.tune(make, config.dambreak_8k(), key="w", cands=cands, steps=2, warmup=0,
                       cache=cache, force=True, finalists=0)
    assert not r3["cached"] and made == [128, 256, 512]


def test_finalists_overrule_a_misleading_first_timing(tmp_path, no_gpu):
    """The first solver built for block 128 is fast, later ones slow (a
    first timing that does not hold); block 256 is steadily faster. The
    first pass ranks 128 first; the finalists' re-timings, two rounds in
    alternating order, pick 256."""
    cands = [dict(block=b, cell_layout=1, jacobi_mode=2, chunk_iters=1, coop_blocks_per_cu=0,
                  graph_iters=0) for b in (128, 256, 512)]
    made = []

    def make(cfg):
        made.append(cfg.block)
        first128 = cfg.block == 128 and made.count(128) == 1
        cost = 0.002 if first128 else (0.004 if cfg.block == 256 else 0.008)
        return FakeSolver(cfg, lambda c: cost)

    r = autotune.tune(make, config.dambreak_8k(), key="w", cands=cands, steps=3, warmup=1,
                      cache=None, finalists=2, rounds=2)
    assert [t["shape"]["block"] for t in sorted(r["table"], key=lambda t: t["ms_per_step"])] == \
        [128, 256, 512]
    assert made == [128, 256, 512, 128, 256, 256, 128]
    assert r["shape"]["block"] == 256
    assert [f["shape"]["block"] for f in r["final"]] == [128, 256]
    assert all(len(f["ms"]) == 2 for f in r["final"])


def test_apply():
    shape = dict(block=512, cell_layout=0, jacobi_mode=1, chunk_iters=0, coop_blocks_per_cu=2,
                 graph_iters=0)
    c = autotune.apply(config.dambreak_1m(), shape)
    assert (c.block, c.cell_layout, c.jacobi_mode, c.coop_blocks_per_cu) == (512, 0, 1, 2)
    assert c.spacing == config.dambreak_1m().spacing
