"""Hydrostatics: once a fluid layer has settled in its box, IISPH's pressure
grows with depth as rho0 g (the steady momentum balance, grad p = rho0 g,
that the pressure solve of IISPH.py:1155-1296 enforces through the
divergence-free density condition). The reference ships no such check; this
pins the whole step (gravity, the pressure solve's scale, the wall's psi, the
pressure force) against physics, on the oracle here and on the HIP path in
the GPU version. Single steps are noisy (the solve stops at the 2-iteration
floor), so the slope is averaged over 100 samples, steps 1,001-2,000."""
import numpy as np
import pytest

from helpers import assert_hydrostatic, hydrostatic_config, hydrostatic_fit
from warpsph_amd import scene

SETTLE, WINDOW, EVERY = 1000, 1000, 10


def test_oracle_settled_layer_is_hydrostatic(orc):
    cfg = hydrostatic_config()
    o = orc.OracleSolver(cfg, scene.build(cfg))
    slopes, surfaces = [], []
    for k in range(1, SETTLE + WINDOW + 1):
        info = o.step()
        if k > SETTLE and k % EVERY == 0:
            a, ys = hydrostatic_fit(cfg, o.x, info["p"])
            slopes.append(a)
            surfaces.append(ys)
    assert np.abs(o.v).max() < 0.5  # settled: no splash left
    assert_hydrostatic(cfg, slopes, surfaces)


@pytest.mark.gpu
def test_hip_settled_layer_is_hydrostatic():
    from warpsph_amd.solver import IISPH
    print("[gpu] hydrostatic layer, 2000 steps", flush=True)
    cfg = hydrostatic_config()
    s = IISPH(stage_path=None, config=cfg)
    slopes, surfaces = [], []
    for k in range(1, SETTLE + WINDOW + 1):
        s.step()
        if k > SETTLE and k % EVERY == 0:
            a, ys = hydrostatic_fit(cfg, s.x.cpu().numpy(), s.p.cpu().numpy())
            slopes.append(a)
            surfaces.append(ys)
    assert np.abs(s.v.cpu().numpy()).max() < 0.5
    assert_hydrostatic(cfg, slopes, surfaces)
    s.close()
