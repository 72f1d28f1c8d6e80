"""Physical and scene constants under the reference's names (wp_sph/main.py:6-61).

Values equal the reference's for the chosen ``scene_type`` (box and gravity
per scene, main.py:35-53); they come from ``warpsph_amd.config`` so the solver
and this module agree. ``wp_sph.solver.IISPH`` builds the scene that
``scene_type`` names, with models read from ``MODEL_DIR``.
"""
import math
from enum import Enum, auto
from pathlib import Path

from warpsph_amd.config import REFERENCE_SCENES, SolverConfig

MODEL_DIR = Path(__file__).parents[1] / "models"


def model_dir(filename):
    return MODEL_DIR / filename


class SceneType(Enum):
    HOUSE = auto()
    PLANE = auto()
    HAND = auto()
    EMPTY = auto()


scene_type = SceneType.EMPTY
DYNAMIC_SCENE = scene_type in [SceneType.PLANE, SceneType.HAND]

_cfg = SolverConfig(box=REFERENCE_SCENES[scene_type.name]["box"],
                    gravity=REFERENCE_SCENES[scene_type.name]["gravity"])
GRAVITY = _cfg.gravity
DIAMETER = _cfg.spacing
OMEGA = _cfg.omega
ETA = _cfg.eta
RHO_0 = _cfg.rho0
INV_SMALL = _cfg.inv_small
FPS = _cfg.fps
BOX_WIDTH, BOX_HEIGHT, BOX_LENGTH = _cfg.box

SMOOTHING_LENGTH = 2 * DIAMETER
TIME_STEP_MAX = 5e-2 * DIAMETER
FLUID_MASS = RHO_0 * DIAMETER ** 3
SIGMA = 8 / (math.pi * SMOOTHING_LENGTH ** 3)
VIS_MU = _cfg.vis_mu * RHO_0
