"""The oracle's ``boundary_psi`` / ``step`` signatures served by libwsph.so.

The pins of tests/test_oracle_pins.py take an object with these two
functions; passing :class:`HipPath` runs the same hand-built scenes through
the C-ABI (one fresh solver per call, one step) so the closed-form checks
apply to the HIP results themselves. Needs a GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from warpsph_amd import _native as N

FIELDS = {"rho": N.F_RHO, "v_adv": N.F_V_ADV, "rho_adv": N.F_RHO_ADV, "d_ii": N.F_D_II,
          "a_ii": N.F_A_II, "p": N.F_P, "Ap": N.F_AP, "s": N.F_S, "ff_count": N.F_FF_COUNT,
          "fs_count": N.F_FS_COUNT}


class HipPath:
    def __init__(self):
        import torch
        self.torch = torch
        self.lib = N.load()

    def _dev(self, a):
        a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
        return self.torch.from_numpy(a).cuda() if a.shape[0] else None

    def _solver(self, P, wt, gt, xb, vb):
        L = self.lib
        h = C.c_void_p()
        N.check("wsph_create", L.wsph_create(C.byref(P), C.byref(h)))
        wt = np.ascontiguousarray(wt, np.float32)
        gt = np.ascontiguousarray(gt, np.float32)
        N.check("wsph_set_tables", L.wsph_set_tables(h, wt.ctypes.data, wt.shape[0],
                                                     gt.ctypes.data, gt.shape[0]))
        xbt, vbt = self._dev(xb), (None if vb is None else self._dev(vb))
        self.torch.cuda.synchronize()
        nb = 0 if xbt is None else xbt.shape[0]
        N.check("wsph_set_boundary", L.wsph_set_boundary(
            h, None if xbt is None else xbt.data_ptr(), None if vbt is None else vbt.data_ptr(),
            nb))
        return h, (xbt, vbt)

    def _field(self, h, fid, count):
        torch = self.torch
        width = N.FIELD_WIDTH[fid]
        dtype = torch.int32 if fid in N.FIELD_IS_INT else torch.float32
        out = torch.empty((count, width) if width > 1 else (count,), dtype=dtype, device="cuda")
        if count:
            torch.cuda.synchronize()
            N.check("wsph_get_field", self.lib.wsph_get_field(h, fid, out.data_ptr()))
        return out.cpu().numpy()

    def boundary_psi(self, P, wt, xb):
        h, keep = self._solver(P, wt, wt, xb, None)
        try:
            return self._field(h, N.F_PSI, np.asarray(xb).reshape(-1, 3).shape[0])
        finally:
            self.lib.wsph_destroy(h)

    def step(self, P, wt, gt, x, v, xb, vb, psi, p0=None, residual_mode=0, want_fields=True):
        if p0 is not None:
            raise ValueError("HipPath.step runs the zero start only")
        n = np.asarray(x).reshape(-1, 3).shape[0]
        h, keep = self._solver(P, wt, gt, xb, vb)
        try:
            xt, vt = self._dev(x), self._dev(v)
            self.torch.cuda.synchronize()
            N.check("wsph_set_fluid", self.lib.wsph_set_fluid(h, xt.data_ptr(), vt.data_ptr(), n))
            st = N.WsphStepStats()
            N.check("wsph_step", self.lib.wsph_step(h, C.byref(st)))
            info = {k: self._field(h, f, n) for k, f in FIELDS.items()}
            info.update(iterations=st.iterations, converged=bool(st.converged),
                        residual=st.residual, v_max=np.float32(st.v_max),
                        max_ff=st.max_ff_count, max_fs=st.max_fs_count)
            return self._field(h, N.F_X, n), self._field(h, N.F_V, n), info
        finally:
            self.lib.wsph_destroy(h)
