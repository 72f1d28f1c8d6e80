"""Property tests of the oracle's neighbour search (IISPH.py:52-163) against
brute force on hypothesis-generated inputs: random clouds, coordinates of
either sign, pairs within a few float32 ulps of h, and hash dims from 1 to
40 per axis (small dims alias far cells into one bucket; the query range is
clamped to dims cells per axis, so each bucket is visited once)."""
import numpy as np
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from helpers import brute_force_lists
from neighbor_cases import point_set

H = 0.04
SETTINGS = settings(max_examples=60, deadline=None, derandomize=True, database=None,
                    suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])
DIMS = st.tuples(*[st.integers(1, 40)] * 3)


@SETTINGS
@given(case=point_set(H), dims=DIMS)
def test_fluid_sets_equal_brute_force(orc, case, dims):
    x, _ = case
    cnt, idx, _ = orc.neighbors(x, x, H, dims, True, cap=256)
    bf = brute_force_lists(x, x, H, True)
    for i in range(x.shape[0]):
        got = idx[i, :cnt[i]].tolist()
        assert len(got) == len(set(got)), i
        assert set(got) == bf[i], i
    # the fluid-fluid relation is symmetric
    pairs = {(i, int(j)) for i in range(x.shape[0]) for j in idx[i, :cnt[i]]}
    assert all((j, i) in pairs for (i, j) in pairs)


@SETTINGS
@given(fluid=point_set(H), boundary=point_set(H), dims=DIMS)
def test_boundary_sets_equal_brute_force(orc, fluid, boundary, dims):
    x, _ = fluid
    xb, _ = boundary
    # half the boundary points moved next to fluid points: cross pairs near h
    m = min(len(x), len(xb) // 2)
    xb = xb.copy()
    if m > 0:
        xb[:m] = (x[:m].astype(np.float64) + (xb[:m] - xb[:m].mean(0)) * 1e-3
                  + np.float32(H) * np.array([1.0, 0.0, 0.0])).astype(np.float32)
    cnt, idx, _ = orc.neighbors(x, xb, H, dims, False, cap=256)
    bf = brute_force_lists(x, xb, H, False)
    for i in range(x.shape[0]):
        got = idx[i, :cnt[i]].tolist()
        assert len(got) == len(set(got)), i
        assert set(got) == bf[i], i
