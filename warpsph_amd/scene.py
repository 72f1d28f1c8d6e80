"""Scene generators.

``reference_box`` restates IISPH.init_particles (wp_sph/solver/IISPH.py:857-1000)
for the reference's EMPTY scene: a lattice shell of wall particles built by a
triple loop in x-major order (IISPH.py:872-912) and a fluid block from
``np.arange`` + ``meshgrid(indexing="ij")`` (IISPH.py:941-949), both cast from
float64 to float32 (IISPH.py:990-994). ``dam_break`` builds the lattice block
used by the benchmark workloads in the same order and precision.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .config import SolverConfig


def _c32(a: np.ndarray) -> np.ndarray:
    """C-contiguous float32 copy (device handoff reads raw row-major memory)."""
    return np.ascontiguousarray(a, dtype=np.float32)


@dataclass
class Scene:
    x: np.ndarray            # (n, 3) float32 fluid positions
    xb: np.ndarray           # (nb, 3) float32 boundary positions
    vb: np.ndarray           # (nb, 3) float32 boundary velocities
    boundary_hide_n: int     # wall particles come first (IISPH.py:927)
    bounds_lo: np.ndarray    # float64 lower corner of everything
    bounds_hi: np.ndarray    # float64 upper corner of everything

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def nb(self) -> int:
        return int(self.xb.shape[0])


def box_shell(box: Sequence[float], spacing: float, layers: int) -> np.ndarray:
    """Wall particles of ``layers`` lattice layers around the box interior
    [0, int(B/d) - 1]^3, enumerated x outer, y, z inner (IISPH.py:872-912).
    Returns float64 positions idx * spacing."""
    ni = [int(b / spacing) for b in box]
    rng = [np.arange(-layers, k + layers) for k in ni]
    ix, iy, iz = np.meshgrid(*rng, indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    outside = ((ix < 0) | (ix > ni[0] - 1) | (iy < 0) | (iy > ni[1] - 1)
               | (iz < 0) | (iz > ni[2] - 1))
    idx = np.stack([ix[outside], iy[outside], iz[outside]], axis=1)
    return idx.astype(np.float64) * spacing


def _bounds(*arrs: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    pts = [a for a in arrs if a.size]
    lo = np.min([a.min(axis=0) for a in pts], axis=0).astype(np.float64)
    hi = np.max([a.max(axis=0) for a in pts], axis=0).astype(np.float64)
    return lo, hi


def reference_box(cfg: SolverConfig) -> Scene:
    """The reference's scenes (IISPH.py:751-766, 857-1000): the wall shell, the
    fluid block, and the model of ``cfg.extra["model"]`` if any."""
    bw, bh, bl = cfg.box
    min_point = cfg.fluid_min or (bw * 0.02, bh * 0.02, bl * 0.02)
    max_point = cfg.fluid_max or (bw * 0.98, bh * 0.4, bl * 0.55)
    for a in range(3):
        if not (0 < min_point[a] < max_point[a] < cfg.box[a]):
            raise ValueError("Invalid range for rectangle")  # IISPH.py:931-939
    xb = box_shell(cfg.box, cfg.spacing, cfg.boundary_layers)
    hide = xb.shape[0]
    axes = [np.arange(min_point[a], max_point[a], cfg.spacing) for a in range(3)]
    xx, yy, zz = np.meshgrid(*axes, indexing="ij")
    x = np.stack([xx.ravel(), yy.ravel(), zz.ravel()]).T
    if cfg.extra.get("model"):  # the scene's model after the walls (IISPH.py:951-989)
        xb = np.concatenate([xb, model_points(cfg.extra["model"], cfg.spacing)])
    x32 = _c32(x)
    xb32 = _c32(xb)
    lo, hi = _bounds(x32, xb32)
    return Scene(x32, xb32, np.zeros_like(xb32), hide, lo, hi)


def lattice_block(cells: Sequence[int], spacing: float,
                  origin: Sequence[float] = (0.0, 0.0, 0.0),
                  jitter: float = 0.0, seed: int = 0) -> np.ndarray:
    """(i, j, k) * spacing + origin in x-major ("ij") order, plus a uniform
    jitter of +-jitter*spacing per coordinate from ``default_rng(seed)``;
    float64."""
    axes = [np.arange(c, dtype=np.float64) * spacing + o for c, o in zip(cells, origin)]
    xx, yy, zz = np.meshgrid(*axes, indexing="ij")
    x = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1)
    if jitter:
        rng = np.random.default_rng(seed)
        x = x + rng.uniform(-jitter, jitter, x.shape) * spacing
    return x


def dam_break(cfg: SolverConfig) -> Scene:
    """Fluid lattice block at ``fluid_origin`` in a box of ``boundary_layers``
    wall layers."""
    xb = box_shell(cfg.box, cfg.spacing, cfg.boundary_layers)
    x = lattice_block(cfg.fluid_cells, cfg.spacing, cfg.fluid_origin, cfg.jitter, cfg.seed)
    x32 = _c32(x)
    xb32 = _c32(xb)
    lo, hi = _bounds(x32, xb32)
    return Scene(x32, xb32, np.zeros_like(xb32), xb32.shape[0], lo, hi)


def model_points(spec: dict, spacing: float) -> np.ndarray:
    """Boundary particles of a model, float64 [m, 3] (the reference's
    ``load_model``, model.py:50-60, and its callers at IISPH.py:951-985).
    ``spec["kind"] == "lattice"``: a solid lattice block of ``cells`` at
    ``origin`` (x-major order, like the mesh samples' meshgrid order);
    ``"mesh"``: the lattice points inside ``spec["mesh"]`` (an OBJ path or a
    procedural mesh, see ``warpsph_amd.mesh``) at ``scale``, rotated by
    ``rot`` and moved to ``pos`` (sampled on the GPU)."""
    kind = spec.get("kind", "lattice")
    d = float(spec.get("spacing", spacing))
    if kind == "lattice":
        return lattice_block(spec["cells"], d, spec.get("origin", (0.0, 0.0, 0.0)))
    if kind == "mesh":
        from . import mesh
        return mesh.load_model(spec["mesh"], spec.get("scale", 1.0),
                               np.asarray(spec.get("pos", (0.0, 0.0, 0.0)), np.float64),
                               np.asarray(spec.get("rot", np.eye(3)), np.float64), d)
    raise ValueError(f"unknown model kind {kind!r}")


def dam_break_model(cfg: SolverConfig) -> Scene:
    """``dam_break`` plus a model appended after the wall particles
    (IISPH.py:951-989: walls first, so boundary_hide_n counts the walls);
    ``cfg.extra["model"]`` describes it (see ``model_points``) and
    ``cfg.rigid_v`` moves it."""
    base = dam_break(cfg)
    xm = _c32(model_points(cfg.extra["model"], cfg.spacing))
    xb32 = _c32(np.concatenate([base.xb, xm]))
    lo, hi = _bounds(base.x, xb32)
    return Scene(base.x, xb32, np.zeros_like(xb32), base.xb.shape[0], lo, hi)


def poisson_surface(shape: str, params: Sequence[float], radius: float, seed: int) -> np.ndarray:
    """Poisson-disk samples (float32 [n, 3]) of a plane rectangle
    (x0, z0, x1, z1, y) or a torus (cx, cy, cz, R, r), axis along y, with
    minimum distance ``radius`` (wsph_poisson_surface; host code)."""
    import ctypes as C
    from . import _native as N
    lib = N.load()
    kind = {"plane": 0, "torus": 1}[shape]
    prm = (C.c_double * 5)(*[float(v) for v in params])
    cnt = C.c_int()
    N.check("wsph_poisson_surface",
            lib.wsph_poisson_surface(kind, prm, float(radius), int(seed), 0, None, C.byref(cnt)))
    out = np.zeros((cnt.value, 3), dtype=np.float32)
    N.check("wsph_poisson_surface",
            lib.wsph_poisson_surface(kind, prm, float(radius), int(seed), cnt.value,
                                     out.ctypes.data, C.byref(cnt)))
    return out


def torus_scene(cfg: SolverConfig) -> Scene:
    """A fluid lattice block above a torus lying on a ground plane; both
    surfaces Poisson-disk sampled at the fluid spacing (plane from ``seed``,
    torus from ``seed + 1``). No walls: fluid that leaves the plane falls
    freely. extra: torus_R, torus_r, torus_center, plane (x0, z0, x1, z1, y)."""
    ex = cfg.extra
    d = cfg.spacing
    R, r = ex.get("torus_R", 0.3), ex.get("torus_r", 0.1)
    cx, cy, cz = ex.get("torus_center", (0.6, 0.1 + d, 0.6))
    plane = ex.get("plane", (-0.4, -0.4, 1.6, 1.6, 0.0))
    xp = poisson_surface("plane", plane, d, cfg.seed)
    xt = poisson_surface("torus", (cx, cy, cz, R, r), d, cfg.seed + 1)
    x = lattice_block(cfg.fluid_cells, d, cfg.fluid_origin, cfg.jitter, cfg.seed)
    x32 = _c32(x)
    xb32 = _c32(np.concatenate([xp, xt]))
    lo, hi = _bounds(x32, xb32)
    return Scene(x32, xb32, np.zeros_like(xb32), xp.shape[0], lo, hi)


def batch_configs(cfg: SolverConfig) -> List[SolverConfig]:
    """Configurations of the independent dam-break scenes of a batch.

    From ``default_rng(cfg.seed)``, per scene: a fluid count
    exp(U(ln n_min, ln n_max)); box edges U(0.5, 2) m; a fluid column against
    the box corner whose x and z extents are U(0.15, 0.6) of the box, its
    height set by the count (capped at 0.9 of the box, the footprint then
    widened); +-1% jitter with a per-scene seed. Spacing and physics are
    those of ``cfg``.
    """
    ex = cfg.extra
    ns, n_min, n_max = int(ex.get("n_scenes", 256)), ex.get("n_min", 2000), ex.get("n_max", 64000)
    rng = np.random.default_rng(cfg.seed)
    d = cfg.spacing
    out = []
    for k in range(ns):
        n = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        box = tuple(float(v) for v in rng.uniform(0.5, 2.0, 3))
        fx, fz = rng.uniform(0.15, 0.6, 2)
        # the column starts one spacing inside the walls; keep one spacing free
        room = [int((b - 2 * d) / d) for b in box]
        cx = max(2, min(room[0], int(round(fx * box[0] / d))))
        cz = max(2, min(room[2], int(round(fz * box[2] / d))))
        cy_max = max(2, int(0.9 * box[1] / d))
        cy = int(round(n / (cx * cz)))
        if cy > cy_max:
            grow = np.sqrt(n / (cx * cy_max * cz))
            cx = max(2, min(room[0], int(round(cx * grow))))
            cz = max(2, min(room[2], int(round(cz * grow))))
            cy = cy_max
        cy = max(2, min(cy, cy_max))
        out.append(cfg.replace(scene="dam_break", box=box, fluid_cells=(cx, cy, cz),
                               fluid_origin=(d, d, d), jitter=0.01, seed=cfg.seed * 1000 + k,
                               boundary_layers=1, extra={}))
    return out


GENERATORS = {
    "reference_box": reference_box,
    "dam_break": dam_break,
    "torus": torus_scene,
    "dam_break_model": dam_break_model,
}


def build(cfg: SolverConfig) -> Scene:
    try:
        gen = GENERATORS[cfg.scene]
    except KeyError:
        raise ValueError(f"unknown scene {cfg.scene!r}") from None
    return gen(cfg)


def domain_grid_dims(lo: np.ndarray, hi: np.ndarray, h: float, slack_cells: int = 4,
                     max_dim: int = 1024) -> Tuple[int, int, int]:
    """Hash dims that give every occupied cell of [lo, hi] its own slot.

    Cell coordinates are trunc(p / h) (float32), so the occupied range on an
    axis is [trunc(lo/h), trunc(hi/h)]; queries reach one cell further on each
    side. Points outside the range still work through the modulo hash, they
    only share slots.
    """
    hf = np.float32(h)
    inv = np.float32(1.0) / hf
    dims = []
    for a in range(3):
        c0 = int(np.trunc(np.float32(lo[a]) * inv))
        c1 = int(np.trunc(np.float32(hi[a]) * inv))
        d = c1 - c0 + 1 + 2 + 2 * slack_cells
        dims.append(max(3, min(d, max_dim)))
    return tuple(dims)  # type: ignore


def grid_dims_for(cfg: SolverConfig, scene: Scene) -> Tuple[Tuple[int, int, int], Tuple[int, int, int]]:
    """(fluid dims, boundary dims) for the configured grid flavour."""
    if cfg.grid == "reference":
        d = cfg.reference_grid_dims()
        return d, d
    d = domain_grid_dims(scene.bounds_lo, scene.bounds_hi, cfg.h)
    return d, d
