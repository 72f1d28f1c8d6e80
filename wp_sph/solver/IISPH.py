"""wp_sph.solver.IISPH -- import path of the reference class
(wp_sph/solver/IISPH.py:736); the implementation is warpsph_amd.solver.IISPH.
As in the reference, the scene follows ``wp_sph.main.scene_type`` (models from
``wp_sph.main.MODEL_DIR``) unless a ``config`` is passed."""
from warpsph_amd import config as _config
from warpsph_amd.solver import IISPH as _IISPH

from .. import main as _main


class IISPH(_IISPH):
    def __init__(self, stage_path="example_sph.usd", preview=False, verbose=False, **kw):
        if kw.get("config") is None:
            kw["config"] = _config.reference(_main.scene_type.name, models_dir=str(_main.MODEL_DIR))
        super().__init__(stage_path, preview, verbose, **kw)
