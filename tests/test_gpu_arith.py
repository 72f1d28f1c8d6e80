"""The kernels' exact replacements for the compiler's correctly rounded
float32 reciprocal, square root and division (one divisor, many dividends)
agree with the compiler's versions bit for bit on the device: every float32
pattern for the reciprocal, every non-negative float32 for the square root,
and 2^29 quotient pairs covering every significand of the divisor. These are
the operations the reference's kernels leave to float32 `/` and `wp.length`
(IISPH.py:372-373 d_ij, sph_funcs.py:105-118 grad_spline_W, and the XSPH
weight m / rho_j)."""
import ctypes
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_rounding_helpers_match_compiler():
    print("[gpu] wsph_selftest_arith: exhaustive rounding check", flush=True)
    sys.stdout.flush()
    from warpsph_amd import _native as N
    lib = N.load()
    import torch
    torch.cuda.init()
    out = (ctypes.c_ulonglong * 4)()
    N.check("wsph_selftest_arith", lib.wsph_selftest_arith(out))
    assert list(out) == [0, 0, 0, 0], (
        f"mismatches: reciprocal {out[0]}, sqrt {out[1]}, division {out[2]}, pair sqrt {out[3]}")
