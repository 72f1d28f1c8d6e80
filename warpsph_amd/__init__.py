"""MI355X-native IISPH step (gfx950 HIP kernels behind a C-ABI).

Host modules: ``config`` (solver parameters and workload presets), ``scene``
(scene generators), ``tables`` (kernel tables), ``_native`` (ctypes binding of
libwsph.so) and ``solver`` (the wp_sph.solver.IISPH drop-in class).
"""
__all__ = ["config", "scene", "tables"]
