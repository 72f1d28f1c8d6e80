"""ctypes wrapper of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the float32 CPU restatement of the reference IISPH
step that the tests compare libwsph.so with. Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborc.so"

_lib = None


def build(force: bool = False) -> Path:
    """Compile liborc.so with the Makefile beside this file."""
    if force or not LIB.exists():
        subprocess.run(["make", "-B" if force else "-s", "-C", str(HERE), "liborc.so"],
                       check=True, capture_output=True)
    return LIB


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OrcOut(C.Structure):
    _fields_ = [
        ("rho", C.c_void_p), ("v_adv", C.c_void_p), ("rho_adv", C.c_void_p),
        ("d_ii", C.c_void_p), ("a_ii", C.c_void_p), ("p", C.c_void_p),
        ("Ap", C.c_void_p), ("s", C.c_void_p), ("p_carry", C.c_void_p), ("ff_count", C.c_void_p),
        ("fs_count", C.c_void_p), ("res_hist", C.c_void_p), ("res_hist_len", C.c_int),
        ("iterations", C.c_int), ("ap_evals", C.c_int), ("converged", C.c_int),
        ("max_ff", C.c_int), ("max_fs", C.c_int), ("residual", C.c_double),
        ("v_max", C.c_float),
    ]


def _load(path: Path) -> C.CDLL:
    from warpsph_amd._native import WsphParams
    L = C.CDLL(str(path))
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_hash_keys.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_void_p]
    L.orc_neighbors.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float,
                                C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                C.c_void_p]
    L.orc_boundary_psi.argtypes = [C.POINTER(WsphParams), C.c_void_p, C.c_int, C.c_int,
                                   C.c_void_p, C.c_void_p]
    L.orc_step.argtypes = [C.POINTER(WsphParams), C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                           C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                           C.POINTER(OrcOut)]
    return L


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        build()
        _lib = _load(LIB)
    return _lib


@contextlib.contextmanager
def using(path):
    """Route every call of this module through another build of the oracle
    for the duration of the block (the tests build copies with planted
    mistakes to show that the pins of each function catch them)."""
    global _lib
    old = _lib
    _lib = _load(Path(path))
    try:
        yield
    finally:
        _lib = old


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def hash_keys(x, cell_width: float, dims) -> np.ndarray:
    x = _f32(x, (-1, 3))
    d = np.ascontiguousarray(dims, dtype=np.int32)
    keys = np.zeros(x.shape[0], dtype=np.int32)
    lib().orc_hash_keys(_ptr(x), x.shape[0], C.c_float(cell_width), _ptr(d), _ptr(keys))
    return keys


def neighbors(xq, xs, h: float, dims, exclude_self: bool, cap: int = 128):
    """Neighbour lists in the reference's order: (counts[nq], idx[nq, cap],
    dist[nq, cap])."""
    xq = _f32(xq, (-1, 3))
    xs = _f32(xs, (-1, 3))
    d = np.ascontiguousarray(dims, dtype=np.int32)
    cnt = np.zeros(xq.shape[0], dtype=np.int32)
    idx = np.full((xq.shape[0], cap), -1, dtype=np.int32)
    dist = np.zeros((xq.shape[0], cap), dtype=np.float32)
    m = lib().orc_neighbors(_ptr(xq), xq.shape[0], _ptr(xs), xs.shape[0], C.c_float(h),
                            _ptr(d), int(exclude_self), cap, _ptr(cnt), _ptr(idx), _ptr(dist))
    if m < 0:
        raise RuntimeError("orc_neighbors failed")
    if m > cap:
        raise RuntimeError(f"neighbour capacity {cap} < {m}")
    return cnt, idx, dist


def mesh_inside(verts, tris, ax, ay, az) -> np.ndarray:
    """Boolean [nx, ny, nz] containment of the lattice points in the closed
    mesh (orc_mesh_inside: per-point ray parity along +z, top-left rule)."""
    v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tris, np.int32).reshape(-1, 3)
    ax, ay, az = (np.ascontiguousarray(a, np.float64) for a in (ax, ay, az))
    out = np.zeros(ax.size * ay.size * az.size, np.uint8)
    L = lib()
    L.orc_mesh_inside.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                  C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    rc = L.orc_mesh_inside(_ptr(v), v.shape[0], _ptr(t), t.shape[0], _ptr(ax), ax.size, _ptr(ay),
                           ay.size, _ptr(az), az.size, _ptr(out))
    if rc:
        raise ValueError("triangle vertex index out of range")
    return out.astype(bool).reshape(ax.size, ay.size, az.size)


def boundary_psi(params, w_table, xb) -> np.ndarray:
    xb = _f32(xb, (-1, 3))
    wt = _f32(w_table)
    psi = np.zeros(xb.shape[0], dtype=np.float32)
    rc = lib().orc_boundary_psi(C.byref(params), _ptr(wt), wt.shape[0], xb.shape[0], _ptr(xb),
                                _ptr(psi))
    if rc:
        raise RuntimeError("orc_boundary_psi failed")
    return psi


def step(params, w_table, g_table, x, v, xb, vb, psi, p0=None, residual_mode: int = 0,
         want_fields: bool = True):
    """One IISPH step on copies of x, v. Returns (x_new, v_new, info) where
    info holds the per-particle fields and the step scalars."""
    x = _f32(x, (-1, 3)).copy()
    v = _f32(v, (-1, 3)).copy()
    xb = _f32(xb, (-1, 3))
    vb = None if vb is None else _f32(vb, (-1, 3))
    psi = _f32(psi)
    wt, gt = _f32(w_table), _f32(g_table)
    p0a = None if p0 is None else _f32(p0)
    n = x.shape[0]
    out = OrcOut()
    keep = {}
    if want_fields:
        for name, w in (("rho", 1), ("v_adv", 3), ("rho_adv", 1), ("d_ii", 3), ("a_ii", 1),
                        ("p", 1), ("Ap", 1), ("s", 3), ("p_carry", 1)):
            a = np.zeros((n, w) if w > 1 else n, dtype=np.float32)
            keep[name] = a
            setattr(out, name, a.ctypes.data)
        for name in ("ff_count", "fs_count"):
            a = np.zeros(n, dtype=np.int32)
            keep[name] = a
            setattr(out, name, a.ctypes.data)
    hist = np.zeros(max(params.max_iters, params.force_iters, 1), dtype=np.float64)
    out.res_hist = hist.ctypes.data
    out.res_hist_len = hist.shape[0]
    rc = lib().orc_step(C.byref(params), _ptr(wt), wt.shape[0], _ptr(gt), gt.shape[0], n,
                        _ptr(x), _ptr(v), _ptr(p0a), xb.shape[0], _ptr(xb), _ptr(vb), _ptr(psi),
                        int(residual_mode), C.byref(out))
    if rc:
        raise RuntimeError("orc_step failed")
    info = dict(keep)
    info.update(iterations=out.iterations, ap_evals=out.ap_evals, converged=bool(out.converged),
                residual=out.residual, v_max=out.v_max, max_ff=out.max_ff, max_fs=out.max_fs,
                res_hist=hist[:out.iterations].copy())
    return x, v, info


class OracleSolver:
    """Multi-step driver over ``step`` that keeps the host clock like
    IISPH.step does (sim_time += dt; CFL dt update, IISPH.py:1367-1371).

    With ``cfg.rigid_v`` set, the boundary particles from
    ``scene.boundary_hide_n`` on form a moving model: in the step at
    sim_time 0 their velocity becomes rigid_v (set_rigid_v, IISPH.py:652-660,
    launched at :1061-1081, before predict_rho_adv, the first reader of the
    boundary velocity), after each step they advance by v * dt in float32
    (update_rigid_x, IISPH.py:663-671, launched at :1355-1364), and the
    next step's boundary grid is built from the moved positions (``step``
    builds it from xb on every call: DYNAMIC_SCENE, IISPH.py:1415-1416). psi
    keeps its initial values."""

    def __init__(self, cfg, scene, grid_dims=None, rigid_first=None):
        from warpsph_amd import scene as scn
        from warpsph_amd import tables
        from warpsph_amd._native import make_params
        self.cfg = cfg
        gd, bd = grid_dims if grid_dims else scn.grid_dims_for(cfg, scene)
        self.params = make_params(cfg, gd, bd)
        self.wt, self.gt = tables.make_tables(cfg.h, cfg.w_size, cfg.grad_size)
        self.x = scene.x.copy()
        self.v = np.zeros_like(self.x)
        self.xb = _f32(scene.xb, (-1, 3)).copy()
        self.vb = _f32(scene.vb, (-1, 3)).copy()
        self.psi = boundary_psi(self.params, self.wt, self.xb)
        rv = getattr(cfg, "rigid_v", None)
        self.rigid_first = (None if rv is None else scene.boundary_hide_n) \
            if rigid_first is None else rigid_first
        self.rigid_v = np.asarray(rv if rv is not None else (0.0, 0.0, 0.0), dtype=np.float32)
        self.p = np.zeros(self.x.shape[0], dtype=np.float32)
        self.p_carry = self.p.copy()
        self.dt = cfg.dt0
        self.sim_time = 0.0
        self.last = None

    def step(self, p0=None, want_fields=True):
        cfg = self.cfg
        self.params.dt = self.dt
        if p0 is None and cfg.warm_start == 1:
            p0 = self.p_carry
        k = self.rigid_first
        if k is not None and self.sim_time == 0.0:
            self.vb[k:] = self.rigid_v
        self.x, self.v, info = step(self.params, self.wt, self.gt, self.x, self.v, self.xb,
                                    self.vb, self.psi, p0=p0, want_fields=True)
        if k is not None:
            self.xb[k:] = self.xb[k:] + self.vb[k:] * np.float32(self.dt)
        self.p = info["p"]
        self.p_carry = info["p_carry"]
        info["dt"] = self.dt
        self.sim_time += self.dt
        if cfg.dt_mode == 0:
            self.dt = min(cfg.cfl * cfg.spacing / max(float(info["v_max"]), cfg.inv_small),
                          cfg.dt_max)
        info["sim_time"] = self.sim_time
        self.last = info
        return info
