"""Edge cases of the HIP path: no fluid, isolated particles, no boundary,
the neighbour-list writer past its ring and at exact capacities,
neighbour-list capacity overflow (an error, where the reference overruns its
n*60 slots), the reported largest neighbour counts, a batch containing an
empty scene."""
import sys

import numpy as np
import pytest

from helpers import assert_same, oracle_setup
from warpsph_amd import config, scene
from warpsph_amd import _native as N

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _solver(cfg, sc):
    from warpsph_amd.solver import IISPH
    return IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)


def _scene(x, xb):
    x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
    xb = np.ascontiguousarray(xb, np.float32).reshape(-1, 3)
    pts = [a for a in (x, xb) if a.size]
    lo = np.min([a.min(0) for a in pts], 0).astype(np.float64)
    hi = np.max([a.max(0) for a in pts], 0).astype(np.float64)
    return scene.Scene(x, xb, np.zeros_like(xb), xb.shape[0], lo, hi)


def test_no_fluid():
    _announce("edge: no fluid particles")
    cfg = config.dambreak_8k()
    base = scene.build(cfg)
    s = _solver(cfg, _scene(np.zeros((0, 3)), base.xb))
    for _ in range(3):
        s.step()
    st = s.step_stats()
    assert st["iterations"] == 0 and st["converged"]
    assert s.x.shape == (0, 3)
    assert s.counters()["steps"] == 3


@pytest.mark.parametrize("start", [config.WARM_ZERO, config.WARM_PREVIOUS])
def test_isolated_particles_match_oracle(orc, start):
    _announce(f"edge: isolated particles, start {start}")
    cfg = config.dambreak_8k().replace(warm_start=start)
    base = scene.build(cfg)
    # a few particles far from each other and from the walls, one pair within h
    x = np.array([[0.3, 0.5, 0.3], [0.6, 0.6, 0.6], [0.5, 0.2, 0.7], [0.51, 0.2, 0.7]])
    sc = _scene(x, base.xb)
    s = _solver(cfg, sc)
    o = orc.OracleSolver(cfg, sc)
    for _ in range(5):
        s.step()
        info = o.step()
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("v", s.v.cpu().numpy(), o.v)
    assert_same("p", s.p.cpu().numpy(), info["p"])
    assert np.all(np.isfinite(s.x.cpu().numpy()))


def test_no_boundary_free_fall_matches_oracle(orc):
    _announce("edge: no boundary particles")
    cfg = config.dambreak_8k().replace(fluid_cells=(6, 6, 6), warm_start=config.WARM_PREVIOUS)
    x = scene.lattice_block(cfg.fluid_cells, cfg.spacing, (0.3, 0.3, 0.3))
    sc = _scene(x, np.zeros((0, 3)))
    s = _solver(cfg, sc)
    o = orc.OracleSolver(cfg, sc)
    for _ in range(4):
        s.step()
        info = o.step()
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("p", s.p.cpu().numpy(), info["p"])
    assert s.fs_neighbor_num.cpu().numpy().max() == 0


def _lists_match_oracle(orc, cfg, sc, s):
    """Fluid and boundary lists equal the oracle's; returns the fluid counts."""
    _, P, _, _ = oracle_setup(cfg, sc)
    counts = []
    for which, (xs, excl, dims) in enumerate([(sc.x, True, P.grid_dims),
                                              (sc.xb, False, P.bgrid_dims)]):
        cnt_o, idx_o, _ = orc.neighbors(sc.x, xs, cfg.h, tuple(dims), excl, cap=64)
        cnt, idx = s.neighbors(which, cap=64)
        cnt, idx = cnt.cpu().numpy(), idx.cpu().numpy()
        assert_same(f"count[{which}]", cnt, cnt_o)
        for i in range(sc.n):
            assert idx[i, :cnt[i]].tolist() == idx_o[i, :cnt_o[i]].tolist(), (which, i)
        counts.append(cnt_o)
    return counts[0]


def test_lists_past_the_ring(orc):
    """One wave whose lanes have 39 and 0 neighbours: a 40-particle cluster
    (every pair closer than h) and 24 isolated particles. The wave's lowest
    count stays 0, so its flushed list rows stay at 0 and the cluster's lanes
    store their entries 32 onwards straight to the list instead of through
    the 32-row ring; the lists must still equal the oracle's."""
    _announce("edge: neighbour lists past the writer's ring")
    cfg = config.dambreak_8k()
    base = scene.build(cfg)
    rng = np.random.default_rng(5)
    cluster = 0.5 + rng.uniform(-0.01, 0.01, size=(40, 3))  # diagonal 0.035 < h = 0.04
    lone = np.array([[0.1 * k, y, z] for (y, z) in ((0.2, 0.2), (0.2, 0.8), (0.8, 0.2))
                     for k in range(1, 9)])
    sc = _scene(np.concatenate([cluster, lone]), base.xb)
    s = _solver(cfg, sc)
    s.neighbor_search()
    cnt = _lists_match_oracle(orc, cfg, sc, s)
    assert cnt[:40].min() == 39 and cnt[40:].max() == 0


def test_lists_at_their_largest_counts(orc):
    """Capacities at the scene's largest neighbour counts (rounded up to the
    library's multiple of 4): the writer's ring limit then comes from the
    capacity (below the ring's 32 rows for the boundary lists), and the lists
    still equal the oracle's."""
    cfg = config.dambreak_8k()
    sc = scene.build(cfg)
    _, P, _, _ = oracle_setup(cfg, sc)
    mf = int(orc.neighbors(sc.x, sc.x, cfg.h, tuple(P.grid_dims), True, cap=64)[0].max())
    mb = int(orc.neighbors(sc.x, sc.xb, cfg.h, tuple(P.bgrid_dims), False, cap=64)[0].max())
    _announce(f"edge: neighbour lists at capacity {mf}/{mb}")
    assert mb < 32 <= mf
    cfg = cfg.replace(max_ff=-(-mf // 4) * 4, max_fs=-(-mb // 4) * 4)
    s = _solver(cfg, sc)
    s.neighbor_search()
    _lists_match_oracle(orc, cfg, sc, s)


def test_neighbour_capacity_overflow_is_an_error():
    _announce("edge: neighbour capacity overflow")
    cfg = config.dambreak_8k().replace(max_ff=8)
    s = _solver(cfg, scene.build(cfg))
    with pytest.raises(RuntimeError, match="capacity"):
        s.step()
    # a solver with room continues normally on the same scene
    cfg2 = config.dambreak_8k()
    s2 = _solver(cfg2, scene.build(cfg2))
    s2.step()
    assert s2.step_stats()["max_ff_count"] <= cfg2.max_ff


def test_reported_count_maxima_are_each_steps_largest_counts():
    """The list pass records each wave's largest counts in one of 256 slots
    (cleared at every step's start, folded at its end): the step statistics
    equal the largest per-particle counts of that step's lists, also in steps
    where the maximum falls (a slot kept from an earlier step would show)."""
    _announce("edge: reported neighbour-count maxima")
    cfg = config.dambreak_8k()
    s = _solver(cfg, scene.build(cfg))
    seen_fall = False
    prev = None
    for _ in range(40):
        s.step()
        st = s.step_stats()
        ff = int(s.field(N.F_FF_COUNT).max().item())
        fs = int(s.field(N.F_FS_COUNT).max().item())
        assert (st["max_ff_count"], st["max_fs_count"]) == (ff, fs)
        if prev is not None and ff < prev:
            seen_fall = True
        prev = ff
    assert seen_fall


def test_particle_count_limit_is_an_error():
    """More than 2^28 - 1 particles (the gathers' 32-bit byte offsets) is a
    capacity error, raised before the pointer is read."""
    _announce("edge: particle count limit")
    cfg = config.dambreak_8k()
    s = _solver(cfg, scene.build(cfg))
    x = s.x
    for fn in ("wsph_set_fluid", "wsph_set_boundary"):
        rc = getattr(s._lib, fn)(s._h, x.data_ptr(), None, 1 << 28)
        assert rc == N.WSPH_E_CAPACITY, (fn, rc)
        assert b"at most" in s._lib.wsph_last_error()
    s.step()  # the solver is unchanged


def test_batch_with_an_empty_scene():
    from warpsph_amd.batch import BatchIISPH
    from warpsph_amd.solver import IISPH
    _announce("edge: batch with an empty scene")
    cfgs = [c.replace(warm_start=config.WARM_PREVIOUS)
            for c in scene.batch_configs(config.batch_256())[:3]]
    scs = [scene.build(c) for c in cfgs]
    scs[1] = _scene(np.zeros((0, 3)), scs[1].xb)
    b = BatchIISPH(cfgs, scs)
    for _ in range(4):
        b.step()
    st = b.scene_stats()
    assert st["converged"].all()
    s0 = IISPH(stage_path=None, config=cfgs[0], scene=scs[0])
    for _ in range(4):
        s0.step()
    assert_same("scene 0 x", b.split(b.x.cpu().numpy())[0], s0.x.cpu().numpy())


def test_demo_script_runs(tmp_path):
    """scenes/IISPH_demo.py (the reference's demo command line) runs the
    reference scene for 40 steps in a child process and records the fluid
    and the boundary model as USDA time samples."""
    import subprocess
    from pathlib import Path
    _announce("demo script, 40 steps")
    root = Path(__file__).resolve().parents[1]
    stage = tmp_path / "demo.usda"
    r = subprocess.run([sys.executable, str(root / "scenes" / "IISPH_demo.py"), "--num_frames", "40",
                        "--stage_path", str(stage), "--no-compress"],
                       capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    text = stage.read_text()
    assert "fluid" in text and "timeSamples" in text
