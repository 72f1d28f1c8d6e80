"""wp_sph.solver.IISPH -- import path of the reference class
(wp_sph/solver/IISPH.py:736); the implementation is warpsph_amd.solver.IISPH."""
from warpsph_amd.solver import IISPH  # noqa: F401

from ..main import *  # noqa: F401,F403
