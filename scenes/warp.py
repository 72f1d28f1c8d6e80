"""Minimal ``warp`` module for the scene scripts in this folder.

The reference's demo (scenes/IISPH_demo.py:5, :46) imports ``warp`` only to
select a device with ``wp.ScopedDevice(args.device)`` before it builds the
solver. Warp is not used here: ``ScopedDevice`` selects the ROCm device
through torch and nothing else is provided. A script in this folder finds
this file first on ``sys.path``.
"""
from __future__ import annotations

__all__ = ["ScopedDevice"]


class ScopedDevice:
    """Context manager that makes ``device`` ("cuda", "cuda:N" or None for the
    current device) current for the enclosed block."""

    def __init__(self, device=None):
        self.device = device
        self._prev = None

    def __enter__(self):
        import torch
        if self.device is not None and str(self.device).startswith("cuda"):
            self._prev = torch.cuda.current_device()
            torch.cuda.set_device(torch.device(str(self.device)))
        elif self.device not in (None, "cuda"):
            raise RuntimeError(f"device {self.device!r}: warpsph_amd runs on ROCm GPUs only")
        return self

    def __exit__(self, *exc):
        if self._prev is not None:
            import torch
            torch.cuda.set_device(self._prev)
        return False
