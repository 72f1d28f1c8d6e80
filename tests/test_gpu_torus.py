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


def test_torus_4m_full_size():
    """The 4,096,000-particle block falling onto the torus and the plane at
    full size (neighbour queries IISPH.py:52-163 over an unbounded domain,
    no walls), 150 steps with the previous-step start: the block lands on
    the torus around step 100. Every step converges (IISPH.py:1200-1207),
    the neighbour lists stay within their capacities (no WSPH_E_CAPACITY,
    largest counts below max_ff / max_fs), and the device-driven and
    persistent Jacobi drivers give identical states and iterations."""
    from warpsph_amd.solver import IISPH
    steps = 150
    cfg = config.torus_4m().replace(warm_start=config.WARM_PREVIOUS)
    sc = scene.build(cfg)
    assert sc.n == 4096000
    out = {}
    for mode in (config.JACOBI_DEVICE, config.JACOBI_PERSISTENT):
        print(f"[gpu] torus 4M, {steps} steps, jacobi={mode}", flush=True)
        sys.stdout.flush()
        s = IISPH(stage_path=None, config=cfg.replace(jacobi_mode=mode), scene=sc,
                  collect_stats=True)
        its, mff, mfs = [], 0, 0
        for k in range(steps):
            s.step()  # raises on WSPH_E_CAPACITY / WSPH_E_NAN
            st = s.step_stats()
            assert st["converged"] and st["status"] == 0, (k, st)
            its.append(st["iterations"])
            mff, mfs = max(mff, st["max_ff_count"]), max(mfs, st["max_fs_count"])
        c = s.counters()
        assert c["nonconverged"] == 0 and c["status"] == 0
        assert 0 < mff < cfg.max_ff and 0 < mfs < cfg.max_fs, (mff, mfs)
        out[mode] = (its, s.x.cpu().numpy(), s.v.cpu().numpy(), s.p.cpu().numpy())
        print(f"[gpu] torus 4M jacobi={mode}: {sum(its) / steps:.2f} iterations/step, "
              f"largest neighbour counts {mff} fluid, {mfs} boundary", flush=True)
        s.close()
    ref = out[config.JACOBI_DEVICE]
    got = out[config.JACOBI_PERSISTENT]
    assert got[0] == ref[0]
    for name, a, b in zip(("x", "v", "p"), got[1:], ref[1:]):
        assert_same(name, a, b)
