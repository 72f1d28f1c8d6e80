"""Host side of the IISPH step: the ``wp_sph.solver.IISPH`` class.

Mirrors the reference class (wp_sph/solver/IISPH.py:736-1501): same
constructor signature, ``step()``, ``render()``, ``neighbor_search()``,
``average_rho_error``, ``window_closed``, ``raise_error`` and state attributes.
The per-particle work runs in libwsph.so (HIP, gfx950) through ctypes; torch
ROCm tensors carry the device memory handed to and from the library.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N
from . import config as cfgmod
from . import scene as scn
from . import tables


def _torch():
    import torch
    return torch


def _sync_torch() -> None:
    """Finish torch's queued work before the library's stream reads tensors
    torch wrote (the library runs on its own stream)."""
    _torch().cuda.synchronize()


class IISPH:
    """IISPH solver on one GPU (reference: IISPH.py:736).

    ``stage_path``, ``preview`` and ``verbose`` mean what they mean in the
    reference; ``config`` selects a :class:`SolverConfig` (default: the
    reference's own scene, ``config.reference()``) and ``device`` the torch
    device (default ``cuda:0``).
    """

    def __init__(self, stage_path: Optional[str] = "example_sph.usd", preview: bool = False,
                 verbose: bool = False, config: Optional[cfgmod.SolverConfig] = None,
                 device: Optional[str] = None, scene: Optional[scn.Scene] = None,
                 collect_stats: bool = False, predictor=None):
        torch = _torch()
        self.verbose = verbose
        self.collect_stats = collect_stats
        self.cfg = cfg = config if config is not None else cfgmod.reference()
        self.device = torch.device(device or "cuda:0")
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("IISPH needs a ROCm GPU (torch.cuda); there is no CPU path")
        torch.cuda.set_device(self.device)
        # render params (IISPH.py:742-744)
        self.frame_dt = 1.0 / cfg.fps
        self.sim_time = 0.0
        self.last_frame_time = 0.0
        # simulation params (IISPH.py:747-749)
        self.dt = cfg.dt0
        self.inv_dt = 1 / self.dt
        self.boundary_layer = cfg.boundary_layers

        sc = scene if scene is not None else scn.build(cfg)
        self.scene = sc
        self.n = sc.n
        self.boundary_n = sc.nb
        self.boundary_hide_n = sc.boundary_hide_n
        self.grid_dims, self.bgrid_dims = scn.grid_dims_for(cfg, sc)
        self.W_table, self.grad_W_table = tables.make_tables(cfg.h, cfg.w_size, cfg.grad_size)

        self._lib = N.load()
        self._params = N.make_params(cfg, self.grid_dims, self.bgrid_dims)
        h = C.c_void_p()
        N.check("wsph_create", self._lib.wsph_create(C.byref(self._params), C.byref(h)))
        self._h = h
        N.check("wsph_set_tables", self._lib.wsph_set_tables(
            h, self.W_table.ctypes.data, self.W_table.shape[0],
            self.grad_W_table.ctypes.data, self.grad_W_table.shape[0]))
        bx0 = torch.from_numpy(np.ascontiguousarray(sc.xb, np.float32)).to(self.device)
        bv0 = torch.from_numpy(np.ascontiguousarray(sc.vb, np.float32)).to(self.device)
        _sync_torch()
        N.check("wsph_set_boundary", self._lib.wsph_set_boundary(
            h, bx0.data_ptr(), bv0.data_ptr(), self.boundary_n))
        # a moving model (SceneType.PLANE / HAND, main.py:20, IISPH.py:1061-1081)
        if cfg.rigid_v is not None:
            self.set_rigid(self.boundary_hide_n, cfg.rigid_v)
        x0 = torch.from_numpy(np.ascontiguousarray(sc.x, np.float32)).to(self.device)
        _sync_torch()
        N.check("wsph_set_fluid", self._lib.wsph_set_fluid(h, x0.data_ptr(), None, self.n))
        self.predictor = None
        if predictor is not None or cfg.warm_start == cfgmod.WARM_LEARNED:
            self.set_predictor(predictor)
        self.last_stats: Optional[dict] = None
        self._stats = N.WsphStepStats()
        if verbose:
            print(f"Initialized {self.n} fluid and {self.boundary_n} boundary particles")

        # renderer (IISPH.py:823-827): USD text recorder of particle positions
        if stage_path:
            from .usd import UsdPointRecorder
            self.renderer = UsdPointRecorder(stage_path, fps=cfg.fps)
        else:
            self.renderer = None
        # the OpenGL previewer (IISPH.py:829-855) is a display window; this
        # build has none
        self.previewer = None
        if preview:
            print("preview window unavailable: warpsph_amd has no OpenGL previewer")

    # ---- lifecycle ------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.wsph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the step (IISPH.py:1002-1371) ----------------------------------
    def _need_stats(self) -> bool:
        return (self.collect_stats or self.verbose or self.cfg.dt_mode == cfgmod.DT_CFL
                or self.renderer is not None)

    def step(self):
        """Advance one time step (IISPH.step, IISPH.py:1002)."""
        st = self._stats if self._need_stats() else None
        rc = self._lib.wsph_step(self._h, C.byref(st) if st is not None else None)
        if rc != N.WSPH_OK:
            self.raise_error(f"step failed ({rc}): {self._lib.wsph_last_error().decode()}")
        if st is not None:
            self.last_stats = st.as_dict()
            self.sim_time = st.sim_time
            if self.cfg.dt_mode == cfgmod.DT_CFL:
                _, self.dt = self.get_time()
            if self.verbose and not st.converged:
                print(f"Pressure solver did not converge at {self.sim_time:.2f}s")
        else:
            self.sim_time += self.dt
        self.inv_dt = 1 / self.dt

    def neighbor_search(self):
        """Grid build and neighbour lists (IISPH.py:1408-1482)."""
        N.check("wsph_neighbor_search", self._lib.wsph_neighbor_search(self._h))

    def step_stats(self) -> dict:
        """Statistics of the last step (iterations, residual, v_max, ...)."""
        return dict(self.last_stats or {})

    def get_time(self):
        t, dt = C.c_double(), C.c_double()
        N.check("wsph_get_time", self._lib.wsph_get_time(self._h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def counters(self) -> dict:
        out = (C.c_longlong * 8)()
        N.check("wsph_get_counters", self._lib.wsph_get_counters(self._h, out, 8))
        return dict(steps=out[0], iterations=out[1], nonconverged=out[2], ap_evals=out[3],
                    status=out[4], barrier_seen=out[5], persistent_grid=out[6])

    def history(self, max_records: int = 8192) -> list:
        """Per-step statistics kept on the device: iterations, residual (mean
        density error), kinetic energy, mean and free-surface height, v_max."""
        recs = (N.WsphStepRecord * max_records)()
        cnt = C.c_int()
        N.check("wsph_get_history", self._lib.wsph_get_history(self._h, recs, max_records,
                                                               C.byref(cnt)))
        keys = [k for k, _ in N.WsphStepRecord._fields_]
        return [{k: getattr(recs[i], k) for k in keys} for i in range(cnt.value)]

    # ---- learned start ------------------------------------------------------
    def set_predictor(self, net) -> None:
        """Attach a :class:`warpsph_amd.mlp.PressureNet` for the learned start
        mode (the solver keeps a reference; None detaches)."""
        if net is None and self.cfg.warm_start == cfgmod.WARM_LEARNED:
            raise ValueError("the learned start mode needs a predictor network")
        self.predictor = net
        N.check("wsph_set_predictor", self._lib.wsph_set_predictor(
            self._h, None if net is None else net.handle))

    def set_pair_logging(self, on: bool = True) -> None:
        """Keep each step's (features, target) rows for :meth:`pairs`."""
        N.check("wsph_set_pair_logging", self._lib.wsph_set_pair_logging(self._h, int(bool(on))))

    def pairs(self):
        """(features [n, 8], targets [n]) of the last step, as device tensors."""
        torch = _torch()
        f = torch.empty((self.n, 8), dtype=torch.float32, device=self.device)
        t = torch.empty(self.n, dtype=torch.float32, device=self.device)
        _sync_torch()
        N.check("wsph_get_pairs", self._lib.wsph_get_pairs(self._h, f.data_ptr(), t.data_ptr()))
        return f, t

    def set_rigid(self, first: int, v) -> None:
        """Boundary particles from index ``first`` on move rigidly: velocity
        ``v`` from the step at sim_time 0, positions advanced every step, the
        boundary grid rebuilt every step (set_rigid_v / update_rigid_x /
        DYNAMIC_SCENE, IISPH.py:652-671, 1061-1081, 1355-1364, 1415-1416).
        ``first < 0`` makes the boundary static again."""
        vv = (C.c_float * 3)(*[float(a) for a in (v if v is not None else (0.0, 0.0, 0.0))])
        N.check("wsph_set_rigid", self._lib.wsph_set_rigid(self._h, int(first), vv))

    def set_pressure(self, p) -> None:
        """Caller-supplied start pressure for the next step (original order)."""
        torch = _torch()
        p = torch.as_tensor(p, dtype=torch.float32, device=self.device).contiguous()
        self._p_start = p
        _sync_torch()
        N.check("wsph_set_pressure", self._lib.wsph_set_pressure(self._h, p.data_ptr()))

    # ---- fields (original particle order) -------------------------------
    def field(self, fid: int):
        torch = _torch()
        width = N.FIELD_WIDTH[fid]
        count = self.boundary_n if fid in N.BOUNDARY_FIELDS else self.n
        dtype = torch.int32 if fid in N.FIELD_IS_INT else torch.float32
        shape = (count, width) if width > 1 else (count,)
        out = torch.empty(shape, dtype=dtype, device=self.device)
        if count == 0:
            return out
        _sync_torch()
        N.check("wsph_get_field", self._lib.wsph_get_field(self._h, fid, out.data_ptr()))
        return out

    def neighbors(self, which: int = 0, cap: Optional[int] = None):
        """(counts, indices) of fluid (0) or boundary (1) neighbours in the
        reference's order and original indices."""
        torch = _torch()
        cap = cap or (self.cfg.max_ff if which == 0 else self.cfg.max_fs)
        cnt = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        idx = torch.full((self.n, cap), -1, dtype=torch.int32, device=self.device)
        if self.n == 0:
            return cnt, idx
        _sync_torch()
        N.check("wsph_get_neighbors", self._lib.wsph_get_neighbors(
            self._h, which, cnt.data_ptr(), idx.data_ptr(), cap))
        return cnt, idx

    x = property(lambda self: self.field(N.F_X))
    v = property(lambda self: self.field(N.F_V))
    p = property(lambda self: self.field(N.F_P))
    rho = property(lambda self: self.field(N.F_RHO))
    rho_adv = property(lambda self: self.field(N.F_RHO_ADV))
    v_adv = property(lambda self: self.field(N.F_V_ADV))
    term_d_ii = property(lambda self: self.field(N.F_D_II))
    term_a_ii = property(lambda self: self.field(N.F_A_II))
    term_Ap_i = property(lambda self: self.field(N.F_AP))
    sum_d_ij_p_j = property(lambda self: self.field(N.F_S))
    ff_neighbor_num = property(lambda self: self.field(N.F_FF_COUNT))
    fs_neighbor_num = property(lambda self: self.field(N.F_FS_COUNT))
    boundary_phi = property(lambda self: self.field(N.F_PSI))
    boundary_x = property(lambda self: self.field(N.F_BX))
    boundary_v = property(lambda self: self.field(N.F_BV))

    @property
    def average_rho_error(self) -> float:
        """Mean density error after the last Jacobi iteration (IISPH.py:1489-1494)."""
        if self.last_stats is None:
            return 0.0
        return float(self.last_stats["residual"])

    # ---- rendering (IISPH.py:1373-1406) ----------------------------------
    def render(self):
        if self.renderer is None:
            return
        if self.sim_time - self.last_frame_time >= self.frame_dt:
            self.renderer.begin_frame(self.sim_time)
            self.renderer.render_points("fluid", self.x.cpu().numpy(), radius=self.cfg.spacing / 1.6)
            rigid = self.boundary_x[self.boundary_hide_n:].cpu().numpy()
            self.renderer.render_points("boundary", rigid, radius=self.cfg.spacing / 1.4)
            self.renderer.end_frame()
            self.last_frame_time = self.sim_time

    def raise_error(self, message):
        """IISPH.raise_error (IISPH.py:1484-1487)."""
        if self.renderer:
            self.renderer.save()
        raise RuntimeError(message)

    @property
    def window_closed(self) -> bool:
        """IISPH.window_closed (IISPH.py:1496-1501): there is no window."""
        return False
