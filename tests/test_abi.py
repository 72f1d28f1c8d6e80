"""libwsph.so loads and exports every entry point include/wsph.h declares.
(No compute call: this runs without a GPU.)"""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from warpsph_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "wsph.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wsph_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    bound = sorted(name for name, _, _ in N.SIGNATURES)
    assert declared_symbols() == bound


def test_library_exports_every_symbol():
    if not N.LIB_PATH.exists():
        pytest.fail("libwsph.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(N.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_abi_version_and_error_string():
    lib = N.load()
    header = (Path(N.__file__).resolve().parents[1] / "include" / "wsph.h").read_text()
    declared = int(re.search(r"#define WSPH_ABI_VERSION (\d+)", header).group(1))
    assert lib.wsph_abi_version() == declared == N.ABI_VERSION == 2
    assert isinstance(lib.wsph_last_error(), bytes)


def test_loader_refuses_another_abi_version(tmp_path):
    """A stale build (a library of another WSPH_ABI_VERSION, whose entry points
    may take other arguments) fails at load, not at its first call: a stub
    exporting every bound symbol and reporting version 1."""
    src = tmp_path / "stub.c"
    src.write_text("".join(f"int {name}(void) {{ return 0; }}\n"
                           for name, _, _ in N.SIGNATURES if name != "wsph_abi_version")
                   + "int wsph_abi_version(void) { return 1; }\n")
    so = tmp_path / "libstub.so"
    subprocess.run(["gcc", "-shared", "-fPIC", str(src), "-o", str(so)], check=True)
    with pytest.raises(RuntimeError, match="ABI version 1"):
        N.load(so)


def test_create_rejects_bad_params_without_gpu_work():
    lib = N.load()
    from warpsph_amd import config
    p = N.make_params(config.dambreak_8k(), (6, 6, 6), (6, 6, 6))
    p.max_iters = 1000
    h = ctypes.c_void_p()
    assert lib.wsph_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert b"iteration" in lib.wsph_last_error()
    p = N.make_params(config.dambreak_8k(), (0, 6, 6), (6, 6, 6))
    assert lib.wsph_create(ctypes.byref(p), ctypes.byref(h)) == -1


@pytest.mark.parametrize("h", [0.04, 0.02, 0.01, 0.1, 0.0371])
def test_radius_threshold_matches_sqrt_predicate(h):
    """d2 < threshold exactly when float32 sqrt(d2) < h, checked on every
    float32 within 2^16 ulps of the threshold and on random radicands."""
    import numpy as np
    lib = N.load()
    t = np.float32(lib.wsph_radius_sq_threshold(h))
    hf = np.float32(h)
    bits = t.view(np.uint32)
    near = (np.arange(-65536, 65536, dtype=np.int64) + int(bits)).astype(np.uint32).view(np.float32)
    rnd = np.random.default_rng(1).uniform(0, 4 * float(t), 200000).astype(np.float32)
    for d2 in (near, rnd):
        assert np.array_equal(np.sqrt(d2) < hf, d2 < t)


def test_blocks_per_cu_from_kernel_resources():
    """The co-residency figure of wsph_blocks_per_cu (host arithmetic, no GPU),
    on the persistent Jacobi kernel's own resource figures (the gfx950
    build's kernel-resource-usage remarks) and on limits taken one at a time
    (MI355X: 160 KiB LDS per CU, 8 waves per SIMD)."""
    lib = N.load()
    lds_cu, wps = 160 * 1024, 8
    f = lib.wsph_blocks_per_cu
    # an earlier build of the kernel: 62 VGPRs, 90 SGPRs (84 + VCC and the
    # like), 200 B static + 8,000 B of tables, 256-thread blocks: its SGPRs
    # hold it to 7 per CU, the residency a launch at the occupancy query's 8
    # found (1,792 of 2,048 blocks); with no SGPR figure, the bound for the
    # largest SGPR count, 6
    assert f(256, 62, 90, 200 + 8000, lds_cu, wps) == 7
    assert f(256, 62, 0, 200 + 8000, lds_cu, wps) == 6
    # the same VGPRs with few SGPRs: the wave slots, 8
    assert f(256, 62, 21, 200 + 8000, lds_cu, wps) == 8
    # the current kernel: 104-111 VGPRs (4 waves per SIMD), 106 SGPRs (6),
    # 8,648 B static + 8,000 B of tables; 512-thread blocks put 2 waves on each SIMD
    for sg in (106, 0):
        assert f(512, 114, sg, 8648 + 8000, lds_cu, wps) == 2
        assert f(256, 114, sg, 8648 + 8000, lds_cu, wps) == 4
    # SGPRs: rounded up to 16, plus 16, out of 800 per SIMD -- measured 7 and 6
    # waves per SIMD at 90 and 106; 80 -> 96 -> 8; 112 (the most) -> 6
    assert f(64, 16, 90, 0, lds_cu, 32) == 7
    assert f(64, 16, 106, 0, lds_cu, 32) == 6
    assert f(64, 16, 80, 0, lds_cu, 32) == 8
    assert f(64, 16, 112, 0, lds_cu, 32) == 6
    assert f(64, 16, 113, 0, lds_cu, 32) == N.WSPH_E_ARG
    # LDS: 60,000 B rounds up to 60,416 B, two fit in 160 KiB
    assert f(64, 32, 21, 60000, lds_cu, wps) == 2
    # VGPRs allocated in units of 8: 129 -> 136 -> 3 waves per SIMD
    assert f(256, 129, 21, 0, lds_cu, wps) == 3
    # 16 workgroups per CU at most; one-wave blocks counted on every SIMD (a
    # lower bound: they could pack 4 times as densely)
    assert f(64, 16, 21, 0, lds_cu, wps) == 8
    assert f(64, 16, 21, 0, lds_cu, 32) == 16
    # a block of 1024 threads (16 waves) needs 4 waves on each SIMD
    assert f(1024, 64, 21, 0, lds_cu, wps) == 2
    assert f(0, 64, 21, 0, lds_cu, wps) == N.WSPH_E_ARG


def test_mesh_noalloc_rejects_oversized_lattice_before_device_work():
    """wsph_mesh_inside_noalloc sizes its fixed scratch regions from the
    arguments alone: 9 M lattice columns need 108 MB of column counters,
    beyond its 64 MiB array, so it returns WSPH_E_CAPACITY without touching
    the device (the dummy pointers here are never dereferenced)."""
    lib = N.load()
    dummy = ctypes.c_double(0.0)
    out = ctypes.c_ubyte(0)
    tri = ctypes.c_int(0)
    p = ctypes.addressof(dummy)
    rc = lib.wsph_mesh_inside_noalloc(p, 1, ctypes.addressof(tri), 0, p, 3000, p, 3000, p, 1,
                                      ctypes.addressof(out))
    assert rc == N.WSPH_E_CAPACITY
    assert b"64 MiB" in lib.wsph_last_error()
