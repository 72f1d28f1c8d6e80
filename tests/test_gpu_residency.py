"""Blocks resident on one compute unit, measured, against wsph_blocks_per_cu.

tests/residency/residency (built by __graft_entry__.build()) launches many
more blocks of a shape than fit, each recording its CU and its start and end
times while it spins 200 us with no block waiting on another; the largest
number of overlapping blocks on a CU is the residency. The shapes fix a
kernel's VGPRs and SGPRs by register clobbers: the persistent Jacobi driver
sizes its cooperative grid with wsph_blocks_per_cu, so every figure it gives
must be resident at once. On MI355X the runtime's occupancy query leaves the
SGPRs out (90 and 106 SGPRs: 8 and 7 blocks of 256 per CU by the query, 7 and
6 resident)."""
import subprocess
import sys
from pathlib import Path

import pytest

from warpsph_amd import _native as N

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "residency" / "residency"
# SGPRs the compiler allocates for each shape (the highest clobbered register
# plus VCC and the like; no clobber: the kernel's own 21)
SGPRS = {0: 21, 84: 90, 100: 106}
LDS_PER_CU = 160 * 1024
WAVES_PER_SIMD = 8


def _rows():
    assert EXE.exists(), f"{EXE} is built by __graft_entry__.build()"
    print("[gpu] residency probe", flush=True)
    sys.stdout.flush()
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = []
    for line in r.stdout.splitlines():
        f = line.split()
        kv = {f[k]: int(f[k + 1]) for k in range(1, len(f) - 1, 2)}
        kv["name"] = f[0]
        rows.append(kv)
    return rows


def test_measured_residency_matches_blocks_per_cu():
    lib = N.load()
    rows = _rows()
    assert len(rows) == 9
    for r in rows:
        want = lib.wsph_blocks_per_cu(r["threads"], r["vgpr"], SGPRS[r["sgpr_clobber"]],
                                      r["lds"], LDS_PER_CU, WAVES_PER_SIMD)
        assert want > 0, r
        if r["threads"] >= 256:
            # blocks of 4 or more waves: the figure is exact
            assert r["resident_max"] == want, (r, want)
            assert r["resident_min"] >= want - 1, (r, want)
        else:
            # one-wave blocks: the figure counts a slot on every SIMD (a lower
            # bound); they pack one per SIMD
            assert r["resident_max"] == 4 * want, (r, want)
    # with no SGPR figure (the driver's case) the computation never exceeds
    # what was resident
    for r in rows:
        if r["threads"] >= 256:
            assert lib.wsph_blocks_per_cu(r["threads"], r["vgpr"], 0, r["lds"], LDS_PER_CU,
                                          WAVES_PER_SIMD) <= r["resident_max"], r
