"""ctypes binding of libwsph.so (declared in include/wsph.h).

The library is built in-tree by ``warpsph_amd/build.py`` into
``warpsph_amd/libwsph.so``. Loading fails loudly when it is missing: there is
no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .config import SolverConfig

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libwsph.so"

WSPH_OK = 0
WSPH_E_ARG, WSPH_E_HIP, WSPH_E_CAPACITY, WSPH_E_NAN, WSPH_E_BARRIER, WSPH_E_UNSUPPORTED = \
    -1, -2, -3, -4, -5, -6
F_X, F_V, F_P, F_RHO, F_RHO_ADV, F_V_ADV, F_D_II, F_A_II, F_AP, F_S, \
    F_FF_COUNT, F_FS_COUNT, F_PSI, F_CELL_KEY, F_BX, F_BV = range(16)
FIELD_WIDTH = {F_X: 3, F_V: 3, F_P: 1, F_RHO: 1, F_RHO_ADV: 1, F_V_ADV: 3,
               F_D_II: 3, F_A_II: 1, F_AP: 1, F_S: 3, F_FF_COUNT: 1,
               F_FS_COUNT: 1, F_PSI: 1, F_CELL_KEY: 1, F_BX: 3, F_BV: 3}
#: fields with one row per boundary particle
BOUNDARY_FIELDS = {F_PSI, F_BX, F_BV}
FIELD_IS_INT = {F_FF_COUNT, F_FS_COUNT, F_CELL_KEY}


class WsphParams(C.Structure):
    _fields_ = [
        ("spacing", C.c_double), ("h", C.c_double), ("rho0", C.c_double),
        ("mass", C.c_double), ("gravity", C.c_double), ("omega", C.c_double),
        ("eta", C.c_double), ("inv_small", C.c_double), ("vis_mu", C.c_double),
        ("xsph", C.c_double), ("dt", C.c_double), ("dt_max", C.c_double),
        ("cfl", C.c_double),
        ("dt_mode", C.c_int), ("min_iters", C.c_int), ("max_iters", C.c_int),
        ("force_iters", C.c_int), ("warm_start", C.c_int),
        ("grid_dims", C.c_int * 3), ("bgrid_dims", C.c_int * 3),
        ("cell_layout", C.c_int), ("max_ff", C.c_int), ("max_fs", C.c_int),
        ("jacobi_mode", C.c_int), ("block", C.c_int),
        ("coop_blocks_per_cu", C.c_int), ("chunk_iters", C.c_int), ("graph_iters", C.c_int),
    ]


class WsphStepStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("converged", C.c_int), ("ap_evals", C.c_int),
        ("max_ff_count", C.c_int), ("max_fs_count", C.c_int), ("status", C.c_int),
        ("residual", C.c_double), ("dt", C.c_double), ("sim_time", C.c_double),
        ("v_max", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def make_params(cfg: SolverConfig, grid_dims, bgrid_dims) -> WsphParams:
    p = WsphParams()
    p.spacing = cfg.spacing
    p.h = cfg.h
    p.rho0 = cfg.rho0
    p.mass = cfg.mass
    p.gravity = cfg.gravity
    p.omega = cfg.omega
    p.eta = cfg.eta
    p.inv_small = cfg.inv_small
    p.vis_mu = cfg.vis_mu
    p.xsph = cfg.xsph
    p.dt = cfg.dt0
    p.dt_max = cfg.dt_max
    p.cfl = cfg.cfl
    p.dt_mode = cfg.dt_mode
    p.min_iters = cfg.min_iters
    p.max_iters = cfg.max_iters
    p.force_iters = cfg.force_iters
    p.warm_start = cfg.warm_start
    p.grid_dims[:] = list(grid_dims)
    p.bgrid_dims[:] = list(bgrid_dims)
    p.cell_layout = cfg.cell_layout
    p.max_ff = cfg.max_ff
    p.max_fs = cfg.max_fs
    p.jacobi_mode = cfg.jacobi_mode
    p.block = cfg.block
    p.coop_blocks_per_cu = cfg.coop_blocks_per_cu
    p.chunk_iters = cfg.chunk_iters
    p.graph_iters = cfg.graph_iters
    return p


class WsphStepRecord(C.Structure):
    _fields_ = [
        ("step", C.c_int), ("iterations", C.c_int), ("converged", C.c_int),
        ("start_source", C.c_int),
        ("sim_time", C.c_double), ("residual", C.c_double), ("kinetic_energy", C.c_double),
        ("mean_height", C.c_double), ("surface_height", C.c_double), ("v_max", C.c_double),
        ("start_err_learned", C.c_double), ("start_err_carry", C.c_double),
    ]


_P = C.c_void_p
_I = C.c_int

#: (name, restype, argtypes) of every symbol include/wsph.h declares
SIGNATURES = [
    ("wsph_last_error", C.c_char_p, []),
    ("wsph_abi_version", _I, []),
    ("wsph_radius_sq_threshold", C.c_float, [C.c_float]),
    ("wsph_blocks_per_cu", _I, [_I, _I, _I, _I, _I, _I]),
    ("wsph_selftest_arith", _I, [C.POINTER(C.c_ulonglong)]),
    ("wsph_create", _I, [C.POINTER(WsphParams), C.POINTER(_P)]),
    ("wsph_destroy", _I, [_P]),
    ("wsph_set_tables", _I, [_P, _P, _I, _P, _I]),
    ("wsph_set_boundary", _I, [_P, _P, _P, _I]),
    ("wsph_set_fluid", _I, [_P, _P, _P, _I]),
    ("wsph_set_rigid", _I, [_P, _I, _P]),
    ("wsph_mesh_inside", _I, [_P, _I, _P, _I, _P, _I, _P, _I, _P, _I, _P]),
    ("wsph_mesh_inside_noalloc", _I, [_P, _I, _P, _I, _P, _I, _P, _I, _P, _I, _P]),
    ("wsph_mesh_timings", _I, [_P]),
    ("wsph_set_pressure", _I, [_P, _P]),
    ("wsph_step", _I, [_P, C.POINTER(WsphStepStats)]),
    ("wsph_neighbor_search", _I, [_P]),
    ("wsph_predict_advection", _I, [_P]),
    ("wsph_pressure_solve", _I, [_P, C.POINTER(WsphStepStats)]),
    ("wsph_integrate", _I, [_P, C.POINTER(WsphStepStats)]),
    ("wsph_get_field", _I, [_P, _I, _P]),
    ("wsph_get_neighbors", _I, [_P, _I, _P, _P, _I]),
    ("wsph_get_time", _I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("wsph_hash_keys", _I, [_P, _I, C.c_float, C.POINTER(C.c_int * 3), _P]),
    ("wsph_last_timings", _I, [_P, C.POINTER(C.c_double), _I]),
    ("wsph_set_timing", _I, [_P, _I]),
    ("wsph_get_counters", _I, [_P, C.POINTER(C.c_longlong), _I]),
    ("wsph_get_history", _I, [_P, C.POINTER(WsphStepRecord), _I, C.POINTER(C.c_int)]),
    ("wsph_mlp_create", _I, [_I, C.POINTER(_P)]),
    ("wsph_mlp_destroy", _I, [_P]),
    ("wsph_mlp_nparams", _I, []),
    ("wsph_mlp_set_params", _I, [_P, _P]),
    ("wsph_mlp_get_params", _I, [_P, _P]),
    ("wsph_mlp_forward", _I, [_P, _P, _I, _P]),
    ("wsph_mlp_train_step", _I, [_P, _P, _P, _I, C.c_float, C.POINTER(C.c_float), _P]),
    ("wsph_poisson_surface", _I, [_I, _P, C.c_double, C.c_uint64, _I, _P, C.POINTER(C.c_int)]),
    ("wsph_set_scenes", _I, [_P, _I, _P, _P, _P, _P]),
    ("wsph_get_scene_stats", _I, [_P, _P, _P, _P]),
    ("wsph_set_predictor", _I, [_P, _P]),
    ("wsph_set_pair_logging", _I, [_P, _I]),
    ("wsph_get_pairs", _I, [_P, _P, _P]),
    ("wsph_sync", _I, [_P]),
]

#: WSPH_ABI_VERSION of include/wsph.h, the version SIGNATURES describe
ABI_VERSION = 2

_lib = None


class WsphError(RuntimeError):
    """Raised for a non-zero status of a libwsph call."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libwsph.so (built in-tree); raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else Path(os.environ.get("WSPH_LIB", LIB_PATH))
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m warpsph_amd.build`; the solver has no CPU fallback")
    # torch and the library share the HIP runtime, and torch must bring it up
    # first: with the library loaded before torch's first device query, the
    # device was not found on the MI355X host. Querying is harmless without a
    # GPU.
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    got = lib.wsph_abi_version()
    if got != ABI_VERSION:
        raise RuntimeError(
            f"{p} has ABI version {got}, these bindings describe version {ABI_VERSION}: "
            "rebuild it with `python -m warpsph_amd.build`")
    if path is None:
        _lib = lib
    return lib


def check(fn_name: str, code: int) -> int:
    if code != WSPH_OK:
        lib = load()
        msg = lib.wsph_last_error()
        raise WsphError(fn_name, code, msg.decode() if msg else "")
    return code


def call(fn_name: str, *args) -> int:
    lib = load()
    return check(fn_name, getattr(lib, fn_name)(*args))
