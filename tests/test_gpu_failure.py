"""Failure handling of the HIP path.

* A non-finite velocity in the state makes wsph_step return WSPH_E_NAN (the
  integration pass checks every new position and speed; the Python class
  raises RuntimeError like IISPH.raise_error, IISPH.py:1484-1487). The NaN is
  put in a velocity, not a position, so that during the failing step it only
  meets float arithmetic: the cell and neighbour computations see finite
  positions, and the error is reported once the step has ended.
* eta = 0 cannot be met: the Jacobi loop stops at its 101-iteration limit
  (IISPH.py:1200-1207) and the step reports converged = 0, in the step
  statistics, the history record and the counters, with every field still
  bit-identical to the oracle's, under the device-driven and the host loop.
"""
import sys

import numpy as np
import pytest

from helpers import assert_same, oracle_setup
from warpsph_amd import _native as N
from warpsph_amd import config, scene

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _solver(cfg, sc, v0):
    import torch
    from warpsph_amd.solver import IISPH
    s = IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)
    xt = torch.from_numpy(np.ascontiguousarray(sc.x)).cuda()
    vt = torch.from_numpy(np.ascontiguousarray(v0)).cuda()
    torch.cuda.synchronize()
    N.check("wsph_set_fluid", s._lib.wsph_set_fluid(s._h, xt.data_ptr(), vt.data_ptr(), sc.n))
    return s


@pytest.mark.parametrize("jacobi", [config.JACOBI_DEVICE, config.JACOBI_HOST])
def test_nan_velocity_reports_wsph_e_nan(jacobi):
    cfg = config.dambreak_8k().replace(jacobi_mode=jacobi)
    _announce(f"NaN velocity, jacobi={jacobi}")
    sc = scene.build(cfg)
    v0 = np.zeros_like(sc.x)
    v0[4321, 1] = np.nan
    s = _solver(cfg, sc, v0)
    with pytest.raises(RuntimeError, match=r"\(-4\)"):
        s.step()
    assert s.counters()["status"] & 4  # ST_NAN
    s.close()
    # a clean state steps without error
    s = _solver(cfg, sc, np.zeros_like(sc.x))
    s.step()
    assert s.step_stats()["status"] == 0
    s.close()


@pytest.mark.parametrize("jacobi", [config.JACOBI_DEVICE, config.JACOBI_HOST])
def test_eta_zero_stops_at_the_iteration_limit(orc, jacobi):
    cfg = config.dambreak_8k().replace(eta=0.0, jacobi_mode=jacobi)
    _announce(f"eta = 0, jacobi={jacobi}")
    sc = scene.build(cfg)
    s = _solver(cfg, sc, np.zeros_like(sc.x))
    s.step()
    st = s.step_stats()
    assert st["iterations"] == 101 and not st["converged"] and st["residual"] > 0.0, st
    rec = s.history()[-1]
    assert rec["iterations"] == 101 and rec["converged"] == 0
    assert s.counters()["nonconverged"] == 1
    _, P, wt, gt = oracle_setup(cfg, sc)
    psi = orc.boundary_psi(P, wt, sc.xb)
    x_o, v_o, info = orc.step(P, wt, gt, sc.x, np.zeros_like(sc.x), sc.xb, sc.vb, psi)
    assert info["iterations"] == 101 and not info["converged"]
    assert st["residual"] == info["residual"]
    assert_same("x", s.x.cpu().numpy(), x_o)
    assert_same("v", s.v.cpu().numpy(), v_o)
    assert_same("p", s.p.cpu().numpy(), info["p"])
    s.close()
