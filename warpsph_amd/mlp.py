"""Per-particle pressure predictor: 8 -> 64 -> 64 -> 1 with SiLU, BF16 weights
and FP32 accumulation on the matrix cores (libwsph.so, csrc/mlp.hip).

The network predicts the change of pressure over one step from local
features (see ``FEATURES``); the solver's learned start mode uses it in place
of the reference's p = 0 start (wp_sph/solver/IISPH.py:166-174). Training runs
on the GPU through the library's own forward and backward kernels with Adam;
torch only holds the device tensors.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _native as N

N_IN, N_H = 8, 64
N_PARAM = N_H * N_IN + N_H + N_H * N_H + N_H + N_H + 1

#: feature columns written by the solver's feature pass (csrc/learn.hip)
FEATURES = (
    "density error of the advected state, (rho_adv / rho0 - 1) * 100",
    "a_ii / a_ref (a_ref: a_ii of a particle inside a resting lattice)",
    "fluid neighbour count / 32",
    "start pressure carried from the previous step / P_ref",
    "kernel-weighted 1-hop mean of column 0",
    "kernel-weighted 1-hop mean of column 1",
    "kernel-weighted 1-hop mean of column 2",
    "kernel-weighted 1-hop mean of column 3",
)


def _torch():
    import torch
    return torch


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights for the SiLU layers, zero biases, small output layer."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, dtype=np.float32)
    o = 0
    for fan_in, fan_out, scale in ((N_IN, N_H, 1.0), (N_H, N_H, 1.0), (N_H, 1, 0.1)):
        bound = scale * np.sqrt(6.0 / fan_in)
        p[o:o + fan_in * fan_out] = rng.uniform(-bound, bound, fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    assert o == N_PARAM
    return p


class PressureNet:
    """Owns a ``wsph_mlp`` context: parameters, BF16 copies, Adam moments."""

    def __init__(self, max_batch: int = 65536, params=None, seed: int = 0):
        # torch brings up the GPU runtime first: the tensors handed to the
        # network are torch's
        if not _torch().cuda.is_available():
            raise RuntimeError("PressureNet needs a ROCm GPU (torch.cuda); there is no CPU path")
        self._lib = N.load()
        assert self._lib.wsph_mlp_nparams() == N_PARAM
        h = C.c_void_p()
        N.check("wsph_mlp_create", self._lib.wsph_mlp_create(int(max_batch), C.byref(h)))
        self._h = h
        self.max_batch = int(max_batch)
        self.set_params(init_params(seed) if params is None else params)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.wsph_mlp_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        if p.shape[0] != N_PARAM:
            raise ValueError(f"expected {N_PARAM} parameters, got {p.shape[0]}")
        N.check("wsph_mlp_set_params", self._lib.wsph_mlp_set_params(self._h, p.ctypes.data))

    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAM, dtype=np.float32)
        N.check("wsph_mlp_get_params", self._lib.wsph_mlp_get_params(self._h, p.ctypes.data))
        return p

    def save(self, path) -> None:
        np.save(Path(path), self.params())

    @classmethod
    def load(cls, path, max_batch: int = 65536) -> "PressureNet":
        return cls(max_batch=max_batch, params=np.load(Path(path)))

    def forward(self, x):
        """x: float32 device tensor [n, 8] -> [n]."""
        torch = _torch()
        x = x.contiguous()
        if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != N_IN:
            raise ValueError("x must be float32 [n, 8]")
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        torch.cuda.synchronize()
        N.check("wsph_mlp_forward", self._lib.wsph_mlp_forward(
            self._h, x.data_ptr(), x.shape[0], y.data_ptr()))
        return y

    def train_step(self, x, t, lr: float = 1e-3, want_grad: bool = False):
        """One Adam step on 0.5 mean (net(x) - t)^2 (lr <= 0: evaluate only).
        Returns the loss (and the gradient before the update)."""
        torch = _torch()
        x, t = x.contiguous(), t.contiguous()
        n = x.shape[0]
        if x.shape != (n, N_IN) or t.shape != (n,):
            raise ValueError("x must be [n, 8] and t [n]")
        loss = C.c_float()
        g = np.zeros(N_PARAM, dtype=np.float32) if want_grad else None
        torch.cuda.synchronize()
        N.check("wsph_mlp_train_step", self._lib.wsph_mlp_train_step(
            self._h, x.data_ptr(), t.data_ptr(), n, C.c_float(lr), C.byref(loss),
            None if g is None else g.ctypes.data))
        return (loss.value, g) if want_grad else loss.value
