"""The HIP neighbour lists (k_lists through the C-ABI) on hypothesis-generated
scenes: a random cloud with coordinates of either sign and pairs within a
few float32 ulps of h, split at random into fluid and boundary particles, so
fluid-fluid and fluid-boundary pairs both straddle h. Each list must equal the
oracle's, order included (IISPH.py:52-163), under both cell layouts; the
boundary lists also exercise the near-boundary prefilter of the list pass,
which must not drop a neighbour."""
import sys

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from helpers import assert_same, oracle_setup
from neighbor_cases import point_set
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu

H = config.dambreak_8k().h


def _scene(x, xb):
    pts = [a for a in (x, xb) if a.size]
    lo = np.min([a.min(0) for a in pts], 0).astype(np.float64)
    hi = np.max([a.max(0) for a in pts], 0).astype(np.float64)
    return scene.Scene(np.ascontiguousarray(x, np.float32), np.ascontiguousarray(xb, np.float32),
                       np.zeros_like(xb, dtype=np.float32), xb.shape[0], lo, hi)


@settings(max_examples=25, deadline=None, derandomize=True, database=None,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])
@given(case=point_set(H), split_seed=st.integers(0, 2 ** 32 - 1),
       layout=st.sampled_from([config.LAYOUT_MORTON, config.LAYOUT_LINEAR]))
def test_lists_equal_oracle(orc, case, split_seed, layout):
    from warpsph_amd.solver import IISPH
    pts, _ = case
    is_b = np.random.default_rng(split_seed).random(len(pts)) < 0.4
    is_b[0] = False  # at least one fluid particle
    x, xb = pts[~is_b], pts[is_b]
    cfg = config.dambreak_8k().replace(cell_layout=layout)
    sc = _scene(x, xb)
    print(f"[gpu] neighbour properties: {len(x)} fluid, {len(xb)} boundary, layout {layout}",
          flush=True)
    sys.stdout.flush()
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    s.neighbor_search()
    _, P, _, _ = oracle_setup(cfg, sc)
    for which, (xs, excl, dims) in enumerate([(sc.x, True, P.grid_dims),
                                              (sc.xb, False, P.bgrid_dims)]):
        if which == 1 and sc.nb == 0:
            continue
        cnt_o, idx_o, _ = orc.neighbors(sc.x, xs, cfg.h, tuple(dims), excl, cap=128)
        cnt, idx = s.neighbors(which, cap=128)
        cnt, idx = cnt.cpu().numpy(), idx.cpu().numpy()
        assert_same(f"count[{which}]", cnt, cnt_o)
        for i in range(sc.n):
            assert idx[i, :cnt[i]].tolist() == idx_o[i, :cnt_o[i]].tolist(), (which, i)
    s.close()
