"""Bit-exact parity at full size: the 1,000,000-particle dam break (three
steps) and the 15,625,000-particle one at 0.004 m (one step) against the CPU
oracle, every per-particle field of the last step compared. The Jacobi
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


@pytest.mark.parametrize("preset,steps", [("dambreak_1m", 3), ("dambreak_16m", 1)])
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
