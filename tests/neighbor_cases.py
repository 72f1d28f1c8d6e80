"""Hypothesis strategy of awkward neighbour-search inputs: a random cloud,
coordinates of either sign, and pairs placed at distance h give or take a few
float32 ulps, so that the float32 test sqrt((dx*dx + dy*dy) + dz*dz) < h
(IISPH.py:71-72, :128-129) decides them either way."""
import numpy as np
from hypothesis import strategies as st


def _near_h(h: float, k: int) -> np.float32:
    r = np.float32(h)
    step = np.float32(np.inf) if k > 0 else np.float32(-np.inf)
    for _ in range(abs(k)):
        r = np.nextafter(r, step)
    return r


@st.composite
def point_set(draw, h: float, max_cloud: int = 40, max_pairs: int = 12, spread_choices=(0.05, 0.2, 1.0)):
    """(points float32 (n, 3), number of borderline pairs), n >= 1."""
    seed = draw(st.integers(0, 2 ** 32 - 1))
    n_cloud = draw(st.integers(0, max_cloud))
    n_pairs = draw(st.integers(0 if n_cloud else 1, max_pairs))
    origin = np.array(draw(st.tuples(*[st.floats(-3.0, 3.0, allow_nan=False, width=32)] * 3)),
                      np.float64)
    spread = draw(st.sampled_from(spread_choices))
    axis_aligned = draw(st.booleans())
    ulps = draw(st.lists(st.integers(-4, 4), min_size=n_pairs, max_size=n_pairs))
    rng = np.random.default_rng(seed)
    cloud = origin + rng.uniform(-spread, spread, (n_cloud, 3))
    base = origin + rng.uniform(-spread, spread, (n_pairs, 3))
    if axis_aligned:
        dirs = np.zeros((n_pairs, 3))
        dirs[np.arange(n_pairs), rng.integers(0, 3, n_pairs)] = rng.choice([-1.0, 1.0], n_pairs)
    else:
        dirs = rng.normal(size=(n_pairs, 3))
        dirs /= np.maximum(np.linalg.norm(dirs, axis=1, keepdims=True), 1e-12)
    r = np.array([_near_h(h, k) for k in ulps], np.float64).reshape(-1, 1)
    other = base + dirs * r
    x = np.concatenate([cloud, base, other]).astype(np.float32)
    return x, n_pairs
