"""The closed-form pins of tests/test_oracle_pins.py run on the HIP path.

Every hand-built scene of the pins (one to three particles, interior
lattices, the converged 101-iteration solve) goes through libwsph.so via
:class:`hip_path.HipPath`; each step's fields must equal the oracle's bit for
bit, and the pin's closed forms are then checked on the HIP values
themselves (IISPH.py:217-329 predict_v_adv / predict_rho_adv, :521-539
update_p).
"""
import sys

import numpy as np
import pytest

from helpers import assert_same
from hip_path import HipPath
from test_oracle_pins import ORACLE_PINS

pytestmark = pytest.mark.gpu


class Checked:
    """Runs each call on the HIP path and on the oracle, asserts the results
    are identical, and returns the HIP path's."""

    def __init__(self, orc):
        self.orc = orc
        self.hip = HipPath()

    def boundary_psi(self, P, wt, xb):
        got = self.hip.boundary_psi(P, wt, xb)
        assert_same("psi", got, self.orc.boundary_psi(P, wt, xb))
        return got

    def step(self, P, wt, gt, x, v, xb, vb, psi, **kw):
        hx, hv, hi = self.hip.step(P, wt, gt, x, v, xb, vb, psi, **kw)
        ox, ov, oi = self.orc.step(P, wt, gt, x, v, xb, vb, psi, **kw)
        for k in ("rho", "v_adv", "rho_adv", "d_ii", "a_ii", "p", "Ap", "s", "ff_count",
                  "fs_count"):
            assert_same(k, hi[k], oi[k])
        assert_same("x", hx, ox)
        assert_same("v", hv, ov)
        assert hi["iterations"] == oi["iterations"]
        assert hi["residual"] == oi["residual"]
        return hx, hv, hi


@pytest.mark.parametrize("pin", ORACLE_PINS, ids=lambda f: f.__name__)
def test_hip_path_meets_pin(orc, pin):
    print(f"[gpu] pin {pin.__name__}", flush=True)
    sys.stdout.flush()
    pin(Checked(orc))
