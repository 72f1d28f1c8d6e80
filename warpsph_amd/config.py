"""Solver configuration.

The reference freezes its physics as module globals in wp_sph/main.py:25-61
and in IISPH.__init__ (IISPH.py:742-749, 763-766, 771-780). ``SolverConfig``
holds the same quantities explicitly; its defaults are the reference's
values, and named presets describe the benchmark workloads.
"""
from __future__ import annotations

import dataclasses
import math
from pathlib import Path
from dataclasses import dataclass, field
from typing import Optional, Tuple

from . import tables

DT_CFL = 0
DT_FIXED = 1
WARM_ZERO = 0
WARM_PREVIOUS = 1
WARM_EXTERNAL = 2
WARM_LEARNED = 3
JACOBI_HOST = 0
JACOBI_PERSISTENT = 1
JACOBI_DEVICE = 2
LAYOUT_LINEAR = 0
LAYOUT_MORTON = 1


@dataclass
class SolverConfig:
    # physics (wp_sph/main.py:25-31, 57-61)
    spacing: float = 0.05          # DIAMETER
    rho0: float = 1.0e3            # RHO_0
    gravity: float = 9.8           # GRAVITY (acts along -y)
    omega: float = 0.5             # OMEGA
    eta: float = 1.0e-3            # ETA
    inv_small: float = 1.0e-6      # INV_SMALL
    vis_mu: float = 0.0            # VIS_MU (Laplacian viscosity, IISPH.py:273-278)
    xsph: float = 0.0              # XSPH velocity smoothing coefficient
    fps: int = 60                  # FPS
    # time step (IISPH.py:747, 1368-1370)
    dt_mode: int = DT_CFL
    dt: Optional[float] = None     # None -> TIME_STEP_MAX
    cfl: float = 0.4
    # Jacobi loop bounds (IISPH.py:1200-1207)
    min_iters: int = 2
    max_iters: int = 101
    force_iters: int = 0
    warm_start: int = WARM_ZERO
    # kernel tables (sph_funcs.py:43-44)
    w_size: int = tables.W_SIZE
    grad_size: int = tables.GRAD_W_SIZE
    # scene (wp_sph/main.py:48-51, IISPH.py:749, 763-766)
    box: Tuple[float, float, float] = (4.0, 4.0, 4.0)
    boundary_layers: int = 3
    scene: str = "reference_box"   # generator name in warpsph_amd.scene
    fluid_min: Optional[Tuple[float, float, float]] = None
    fluid_max: Optional[Tuple[float, float, float]] = None
    fluid_cells: Optional[Tuple[int, int, int]] = None
    fluid_origin: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    jitter: float = 0.0            # fraction of spacing
    seed: int = 0
    steps: int = 5000
    # device layout
    grid: str = "domain"           # "domain" (covering) or "reference" (int(BOX/h/4))
    cell_layout: int = LAYOUT_MORTON
    max_ff: int = 128
    max_fs: int = 96
    jacobi_mode: int = JACOBI_DEVICE
    block: int = 256
    coop_blocks_per_cu: int = 0
    chunk_iters: int = 0
    graph_iters: int = 0           # >0: replay the step as a captured hipGraph
    # velocity of the moving model (boundary particles after the walls); None: static
    # boundary (SceneType.PLANE / HAND set (2, 0, 0) / (0, 0.1, 0), IISPH.py:1061-1081)
    rigid_v: Optional[Tuple[float, float, float]] = None
    extra: dict = field(default_factory=dict)

    # ---- derived quantities (wp_sph/main.py:57-60) ---------------------
    @property
    def h(self) -> float:
        """SMOOTHING_LENGTH = 2 DIAMETER."""
        return 2 * self.spacing

    @property
    def mass(self) -> float:
        """FLUID_MASS = RHO_0 DIAMETER^3."""
        return self.rho0 * self.spacing ** 3

    @property
    def sigma(self) -> float:
        """SIGMA = 8 / (pi h^3)."""
        return 8 / (math.pi * self.h ** 3)

    @property
    def dt_max(self) -> float:
        """TIME_STEP_MAX = 0.05 DIAMETER."""
        return 5e-2 * self.spacing

    @property
    def dt0(self) -> float:
        return self.dt if self.dt is not None else self.dt_max

    def reference_grid_dims(self) -> Tuple[int, int, int]:
        """Hash dims of the reference, int(BOX / h / 4) (IISPH.py:771-780)."""
        return tuple(int(b / self.h / 4) for b in self.box)  # type: ignore

    def replace(self, **kw) -> "SolverConfig":
        return dataclasses.replace(self, **kw)


#: the reference's scene types (main.py:13-53, IISPH.py:751-766, 951-985): box,
#: gravity, fluid block as fractions of the box, wall layers, model (OBJ file,
#: scale, position, rotation) and the model's rigid velocity
REFERENCE_SCENES = {
    "EMPTY": dict(box=(4.0, 4.0, 4.0), gravity=9.8, fluid=((0.02, 0.02, 0.02), (0.98, 0.4, 0.55)),
                  layers=3, model=None, rigid_v=None),
    "HOUSE": dict(box=(8.0, 10.0, 12.0), gravity=9.8,
                  fluid=((0.7, 0.01, 0.8), (0.99, 0.5, 0.99)), layers=3,
                  model=("house.obj", 2e-2, (3.0, 0.0, 4.0), None), rigid_v=None),
    "PLANE": dict(box=(8.0, 8.0, 8.0), gravity=0.0,
                  fluid=((0.45, 0.05, 0.05), (0.55, 0.95, 0.95)), layers=0,
                  model=("plane.obj", 3e-3, (0.0, 4.0, 4.0),
                         ((1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0))),
                  rigid_v=(2.0, 0.0, 0.0)),
    "HAND": dict(box=(8.0, 50.0, 8.0), gravity=9.8,
                 fluid=((0.48, 0.05, 0.48), (0.52, 0.99, 0.52)), layers=3,
                 model=("hand.obj", 0.3, (4.5, 2.0, 1.0), "hand"), rigid_v=(0.0, 0.1, 0.0)),
}


def _hand_rotation():
    """The HAND model's rotation, 160 degrees about z, computed with the same
    numpy calls as the reference (IISPH.py:973-981)."""
    import numpy as np
    angle = np.radians(160)
    return tuple(tuple(float(v) for v in row) for row in
                 ((np.cos(angle), -np.sin(angle), 0), (np.sin(angle), np.cos(angle), 0), (0, 0, 1)))


def reference(scene_type: str = "EMPTY", models_dir: Optional[str] = None) -> SolverConfig:
    """The reference's scene of the given SceneType name (default EMPTY: 4x4x4 m
    box, d = 0.05, 3 wall layers, fluid block from (0.02,0.02,0.02) to
    (0.98,0.4,0.55) of the box), CFL dt, 8192/65536-sample tables, p0 = 0
    (main.py:19-61, IISPH.py:738-766). HOUSE, PLANE and HAND add their model,
    sampled from ``<models_dir>/<name>.obj`` (default: the repository's
    ``models/`` folder, as main.py's MODEL_DIR); PLANE and HAND move it."""
    try:
        spec = REFERENCE_SCENES[scene_type]
    except KeyError:
        raise ValueError(f"unknown scene type {scene_type!r}; known: {sorted(REFERENCE_SCENES)}") \
            from None
    box = spec["box"]
    lo, hi = spec["fluid"]
    cfg = SolverConfig(box=box, gravity=spec["gravity"], boundary_layers=spec["layers"],
                       fluid_min=tuple(b * f for b, f in zip(box, lo)),
                       fluid_max=tuple(b * f for b, f in zip(box, hi)),
                       rigid_v=spec["rigid_v"])
    if spec["model"] is not None:
        name, scale, pos, rot = spec["model"]
        mdir = Path(models_dir) if models_dir else Path(__file__).resolve().parents[1] / "models"
        model = {"kind": "mesh", "mesh": str(mdir / name), "scale": scale, "pos": pos}
        if rot is not None:
            model["rot"] = _hand_rotation() if rot == "hand" else rot
        cfg = cfg.replace(extra={"model": model})
    return cfg


def dambreak_8k() -> SolverConfig:
    """20x20x20 = 8,000 fluid particles at 0.02 m, block corner at the origin
    of a 1x1x1 m box with one wall layer; g = 9.81, dt = 1e-3 fixed, XSPH 0.01,
    1000-sample tables; 200 steps."""
    return SolverConfig(
        spacing=0.02, gravity=9.81, xsph=0.01, dt_mode=DT_FIXED, dt=1e-3,
        w_size=1000, grad_size=1000, box=(1.0, 1.0, 1.0), boundary_layers=1,
        scene="dam_break", fluid_cells=(20, 20, 20), steps=200)


def dambreak_1m() -> SolverConfig:
    """100x100x100 = 1,000,000 fluid particles at 0.01 m in a 2x1x1 m box with
    one wall layer, positions jittered by +-1% of the spacing from seed 0;
    otherwise as dambreak_8k; 500 steps."""
    return dambreak_8k().replace(
        spacing=0.01, box=(2.0, 1.0, 1.0), fluid_cells=(100, 100, 100),
        jitter=0.01, seed=0, steps=500)


def dambreak_16m() -> SolverConfig:
    """The 1M dam break's geometry at 0.004 m: 250^3 = 15,625,000 fluid
    particles in the same 2x1x1 m box with one wall layer, jittered by +-1% of
    the spacing from seed 0; dt = 4e-4 s fixed (the 1M scene's dt scaled with
    the spacing); 500 steps. A problem-size point: ~18 GB of device memory."""
    return dambreak_1m().replace(spacing=0.004, fluid_cells=(250, 250, 250), dt=4e-4)


def torus_4m() -> SolverConfig:
    """160^3 = 4,096,000 fluid particles at 0.005 m, block corner at
    (0.2, 0.25, 0.2), falling onto a torus (R = 0.3, r = 0.1, centre
    (0.6, 0.105, 0.6), axis vertical) lying on a 2 x 2 m ground plane at
    y = 0; both surfaces Poisson-disk sampled at 0.005 m from seed 1 (plane)
    and 2 (torus); no walls; dt = 1e-3 fixed, XSPH 0.01; 300 steps."""
    return dambreak_8k().replace(
        spacing=0.005, box=(2.0, 1.1, 2.0), boundary_layers=0,
        scene="torus", fluid_cells=(160, 160, 160), fluid_origin=(0.2, 0.25, 0.2), seed=1,
        steps=300, extra={"torus_R": 0.3, "torus_r": 0.1, "torus_center": (0.6, 0.105, 0.6),
                          "plane": (-0.4, -0.4, 1.6, 1.6, 0.0)})


def batch_256() -> SolverConfig:
    """256 independent dam-break scenes, fluid counts log-uniform in
    2k..64k, random block aspect ratios and box sizes 0.5..2 m from seed 2,
    spacing 0.02 m; 300 steps each."""
    return dambreak_8k().replace(scene="batch", seed=2, steps=300,
                                 extra={"n_scenes": 256, "n_min": 2000,
                                        "n_max": 64000})


def learned_1m() -> SolverConfig:
    """dambreak_1m with the learned pressure start value (predictor network,
    compared with the previous step's pressure); 500 steps."""
    return dambreak_1m().replace(warm_start=WARM_LEARNED)


def paddle_8k() -> SolverConfig:
    """dambreak_8k with a paddle of 3 x 10 x 10 boundary particles at
    (0.44, 0, 0.1), just beyond the fluid block's +x face, pushed into the
    fluid at 1 m/s along -x (a moving model, as the reference's PLANE and
    HAND scenes); 200 steps."""
    return dambreak_8k().replace(
        scene="dam_break_model", rigid_v=(-1.0, 0.0, 0.0),
        extra={"model": {"kind": "lattice", "cells": (3, 10, 10), "origin": (0.44, 0.0, 0.1)}})


def paddle_1m() -> SolverConfig:
    """dambreak_1m with a paddle of 3 x 60 x 60 boundary particles at
    (1.02, 0, 0.2), beyond the fluid block's +x face, moving at 1 m/s along
    -x into the collapsing column; 500 steps."""
    return dambreak_1m().replace(
        scene="dam_break_model", rigid_v=(-1.0, 0.0, 0.0),
        extra={"model": {"kind": "lattice", "cells": (3, 60, 60), "origin": (1.02, 0.0, 0.2)}})


def wheel_8k() -> SolverConfig:
    """dambreak_8k with a mesh-sampled model: a torus (R = 0.1, r = 0.04)
    stood upright (rotated 90 degrees about x) at (0.6, 0.16, 0.5), filled
    with boundary particles at the fluid spacing, moving into the fluid at
    0.5 m/s along -x; 200 steps."""
    rot = ((1.0, 0.0, 0.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0))
    return dambreak_8k().replace(
        scene="dam_break_model", rigid_v=(-0.5, 0.0, 0.0),
        extra={"model": {"kind": "mesh", "mesh": {"mesh": "torus", "center": (0.0, 0.0, 0.0),
                                                  "R": 0.1, "r": 0.04},
                         "pos": (0.6, 0.16, 0.5), "rot": rot}})


PRESETS = {
    "reference": reference,
    "dambreak_8k": dambreak_8k,
    "dambreak_1m": dambreak_1m,
    "dambreak_16m": dambreak_16m,
    "torus_4m": torus_4m,
    "batch_256": batch_256,
    "learned_1m": learned_1m,
    "paddle_8k": paddle_8k,
    "paddle_1m": paddle_1m,
    "wheel_8k": wheel_8k,
}


def preset(name: str) -> SolverConfig:
    try:
        return PRESETS[name]()
    except KeyError:
        raise ValueError(f"unknown preset {name!r}; known: {sorted(PRESETS)}") from None
