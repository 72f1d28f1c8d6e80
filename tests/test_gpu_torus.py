"""The torus-and-plane scene (Poisson-disk sampled boundaries, no walls) at
reduced size: boundary psi and three steps bit-identical to the oracle."""
import sys

import pytest

from helpers import assert_same
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu


def small_torus():
    cfg = config.torus_4m()
    return cfg.replace(fluid_cells=(24, 24, 24), fluid_origin=(0.55, 0.22, 0.55),
                       warm_start=config.WARM_PREVIOUS,
                       extra=dict(cfg.extra, plane=(0.2, 0.2, 1.0, 1.0, 0.0)))


def test_torus_scene_matches_oracle(orc):
    from warpsph_amd.solver import IISPH
    cfg = small_torus()
    print("[gpu] torus scene 24^3 vs oracle", flush=True)
    sys.stdout.flush()
    sc = scene.build(cfg)
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    o = orc.OracleSolver(cfg, sc)
    assert_same("psi", s.boundary_phi.cpu().numpy(), o.psi)
    for k in range(3):
        s.step()
        info = o.step()
        assert s.step_stats()["iterations"] == info["iterations"], k
    assert_same("x", s.x.cpu().numpy(), o.x)
    assert_same("v", s.v.cpu().numpy(), o.v)
    assert_same("p", s.p.cpu().numpy(), info["p"])
    assert_same("rho", s.rho.cpu().numpy(), info["rho"])
