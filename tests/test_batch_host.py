"""Host-side checks of the batched solver (no GPU work)."""
import ctypes

import numpy as np
import pytest

from warpsph_amd import _native as N
from warpsph_amd import config, scene


def test_batch_configs_workload():
    cs = scene.batch_configs(config.batch_256())
    assert len(cs) == 256
    n = np.array([c.fluid_cells[0] * c.fluid_cells[1] * c.fluid_cells[2] for c in cs])
    # log-uniform counts in [2k, 64k] (rounding of the column's cells aside)
    assert n.min() >= 1500 and n.max() <= 70000
    assert 15000 < n.mean() < 22000
    for c in cs:
        assert all(0.5 <= b <= 2.0 for b in c.box)
        # the fluid column fits inside the walls
        for a in range(3):
            assert (c.fluid_cells[a] + 1) * c.spacing < c.box[a]
    # deterministic from the seed
    assert scene.batch_configs(config.batch_256())[17] == cs[17]


def test_batch_rejects_mixed_physics():
    from warpsph_amd.batch import BatchIISPH
    cs = scene.batch_configs(config.batch_256())[:2]
    with pytest.raises(ValueError, match="spacing"):
        BatchIISPH([cs[0], cs[1].replace(spacing=0.03)])


def test_set_scenes_argument_checks():
    lib = N.load()
    cfg = config.dambreak_8k()
    p = N.make_params(cfg, (8, 8, 8), (8, 8, 8))
    p.dt_mode = config.DT_CFL
    h = ctypes.c_void_p()
    if lib.wsph_create(ctypes.byref(p), ctypes.byref(h)) != 0:
        pytest.skip("solver creation needs a GPU")
    counts = np.array([10, 10], np.int32)
    dims = np.full(6, 8, np.int32)
    assert lib.wsph_set_scenes(h, 2, counts.ctypes.data, counts.ctypes.data, dims.ctypes.data,
                               dims.ctypes.data) == -6
    lib.wsph_destroy(h)
