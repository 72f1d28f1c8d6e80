"""Many independent IISPH scenes solved together on one GPU.

The scenes' particles are concatenated (scene order) into one solver; each
scene keeps its own hash grids, so no particle sees another scene's, and each
scene runs its own Jacobi stop test (wsph_set_scenes). Every scene therefore
evolves exactly as it would in a solver of its own, while one launch per pass
covers all of them. The scenes must share the physics and the kernel tables
(spacing, rho0, dt, table sizes), the start mode and the Jacobi driver.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N
from . import config as cfgmod
from . import scene as scn
from . import tables

#: fields that must agree between the scenes of a batch
_SHARED = ("spacing", "rho0", "gravity", "omega", "eta", "inv_small", "vis_mu", "xsph",
           "dt_mode", "dt", "cfl", "min_iters", "max_iters", "force_iters", "warm_start",
           "w_size", "grad_size", "cell_layout", "max_ff", "max_fs", "jacobi_mode", "block",
           "coop_blocks_per_cu", "chunk_iters")


def _torch():
    import torch
    return torch


class BatchIISPH:
    """Solver over a list of scene configurations (one scene each)."""

    def __init__(self, cfgs: Sequence[cfgmod.SolverConfig],
                 scenes: Optional[Sequence[scn.Scene]] = None, device: Optional[str] = None,
                 predictor=None):
        torch = _torch()
        if not cfgs:
            raise ValueError("empty batch")
        base = cfgs[0]
        if any(c.rigid_v is not None for c in cfgs):
            raise ValueError("rigid boundary motion is not available with batched scenes")
        for k, c in enumerate(cfgs):
            for f in _SHARED:
                if getattr(c, f) != getattr(base, f):
                    raise ValueError(f"scene {k}: {f} differs from scene 0")
        self.device = torch.device(device or "cuda:0")
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("BatchIISPH needs a ROCm GPU (torch.cuda); there is no CPU path")
        torch.cuda.set_device(self.device)
        self.cfg = base
        self.cfgs = list(cfgs)
        self.scenes: List[scn.Scene] = list(scenes) if scenes is not None else [
            scn.build(c) for c in cfgs]
        self.nscenes = len(self.scenes)
        self.counts = np.array([s.n for s in self.scenes], dtype=np.int32)
        self.bcounts = np.array([s.nb for s in self.scenes], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        dims = [scn.grid_dims_for(c, s) for c, s in zip(self.cfgs, self.scenes)]
        gd = np.ascontiguousarray([d[0] for d in dims], dtype=np.int32)
        bd = np.ascontiguousarray([d[1] for d in dims], dtype=np.int32)
        self.n = int(self.counts.sum())
        self.boundary_n = int(self.bcounts.sum())
        self.W_table, self.grad_W_table = tables.make_tables(base.h, base.w_size, base.grad_size)
        self._lib = N.load()
        self._params = N.make_params(base, tuple(gd[0]), tuple(bd[0]))
        h = C.c_void_p()
        N.check("wsph_create", self._lib.wsph_create(C.byref(self._params), C.byref(h)))
        self._h = h
        N.check("wsph_set_scenes", self._lib.wsph_set_scenes(
            h, self.nscenes, self.counts.ctypes.data, self.bcounts.ctypes.data, gd.ctypes.data,
            bd.ctypes.data))
        N.check("wsph_set_tables", self._lib.wsph_set_tables(
            h, self.W_table.ctypes.data, self.W_table.shape[0],
            self.grad_W_table.ctypes.data, self.grad_W_table.shape[0]))
        xb = np.concatenate([s.xb for s in self.scenes]).astype(np.float32)
        vb = np.concatenate([s.vb for s in self.scenes]).astype(np.float32)
        x = np.concatenate([s.x for s in self.scenes]).astype(np.float32)
        self.boundary_x = torch.from_numpy(np.ascontiguousarray(xb)).to(self.device)
        self.boundary_v = torch.from_numpy(np.ascontiguousarray(vb)).to(self.device)
        x0 = torch.from_numpy(np.ascontiguousarray(x)).to(self.device)
        torch.cuda.synchronize()
        N.check("wsph_set_boundary", self._lib.wsph_set_boundary(
            h, self.boundary_x.data_ptr(), self.boundary_v.data_ptr(), self.boundary_n))
        N.check("wsph_set_fluid", self._lib.wsph_set_fluid(h, x0.data_ptr(), None, self.n))
        if predictor is not None or base.warm_start == cfgmod.WARM_LEARNED:
            if predictor is None:
                raise ValueError("the learned start mode needs a predictor network")
            self.predictor = predictor
            N.check("wsph_set_predictor", self._lib.wsph_set_predictor(h, predictor.handle))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.wsph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, stats: bool = False):
        st = N.WsphStepStats() if stats else None
        N.check("wsph_step", self._lib.wsph_step(self._h, C.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def scene_stats(self) -> dict:
        """Per-scene Jacobi iterations, convergence and residual of the last step."""
        it = np.zeros(self.nscenes, np.int32)
        cv = np.zeros(self.nscenes, np.int32)
        res = np.zeros(self.nscenes, np.float64)
        N.check("wsph_get_scene_stats", self._lib.wsph_get_scene_stats(
            self._h, it.ctypes.data, cv.ctypes.data, res.ctypes.data))
        return dict(iterations=it, converged=cv.astype(bool), residual=res)

    def counters(self) -> dict:
        out = (C.c_longlong * 8)()
        N.check("wsph_get_counters", self._lib.wsph_get_counters(self._h, out, 8))
        return dict(steps=out[0], iterations=out[1], nonconverged=out[2], ap_evals=out[3],
                    status=out[4], barrier_seen=out[5], persistent_grid=out[6])

    def field(self, fid: int):
        """A field of all scenes (concatenated scene order)."""
        torch = _torch()
        width = N.FIELD_WIDTH[fid]
        count = self.boundary_n if fid in N.BOUNDARY_FIELDS else self.n
        dtype = torch.int32 if fid in N.FIELD_IS_INT else torch.float32
        out = torch.empty((count, width) if width > 1 else (count,), dtype=dtype,
                          device=self.device)
        if count == 0:
            return out
        torch.cuda.synchronize()
        N.check("wsph_get_field", self._lib.wsph_get_field(self._h, fid, out.data_ptr()))
        return out

    def split(self, t) -> list:
        """Per-scene views of a fluid field."""
        return [t[self.offsets[k]:self.offsets[k + 1]] for k in range(self.nscenes)]

    x = property(lambda self: self.field(N.F_X))
    v = property(lambda self: self.field(N.F_V))
    p = property(lambda self: self.field(N.F_P))
    rho = property(lambda self: self.field(N.F_RHO))
