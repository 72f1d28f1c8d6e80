"""Mesh-sampled boundary models.

The reference fills its models (house, plane, hand) with boundary particles
by ``_sample_points`` / ``load_model`` (wp_sph/solver/model.py:14-60): scale
the mesh's vertices, lay an ``np.arange`` lattice of the given spacing over
the scaled bounds (``meshgrid(indexing="ij")``, x outermost), keep the points
``trimesh.Trimesh.contains`` reports inside, then rotate (``points @ rot``)
and translate. This module keeps those steps on the host with the same numpy
expressions and runs the containment test on the GPU
(``wsph_mesh_inside``). Meshes come from OBJ files (``read_obj``) or from the
procedural generators below (the reference's own OBJ assets are not part of
its repository). There is no joblib cache: the GPU test takes milliseconds.
"""
from __future__ import annotations

from pathlib import Path
from typing import Sequence, Tuple, Union

import numpy as np

Mesh = Tuple[np.ndarray, np.ndarray]  # (vertices float64 [nv, 3], triangles int32 [nt, 3])


# ---- meshes -----------------------------------------------------------------
def read_obj(path: Union[str, Path]) -> Mesh:
    """Vertices and triangles of a Wavefront OBJ file (``v`` and ``f``
    records; polygons split into fans; 1-based or negative indices; texture
    and normal indices ignored)."""
    verts, tris = [], []
    for line in Path(path).read_text().splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(a) for a in parts[1:4]])
        elif parts[0] == "f":
            idx = []
            for tok in parts[1:]:
                k = int(tok.split("/")[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)
            for a in range(1, len(idx) - 1):
                tris.append([idx[0], idx[a], idx[a + 1]])
    return np.asarray(verts, np.float64).reshape(-1, 3), np.asarray(tris, np.int32).reshape(-1, 3)


def box_mesh(lo: Sequence[float], hi: Sequence[float]) -> Mesh:
    """Closed axis-aligned box, 12 triangles, outward counter-clockwise."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    v = np.array([[(hi if (k >> a) & 1 else lo)[a] for a in range(3)] for k in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = [(q[0], q[1], q[2]) for q in quads] + [(q[0], q[2], q[3]) for q in quads]
    return v, np.asarray(t, np.int32)


def octahedron(center: Sequence[float], r: float) -> Mesh:
    """Regular octahedron |x| + |y| + |z| <= r around ``center``."""
    c = np.asarray(center, np.float64)
    v = np.array([[r, 0, 0], [-r, 0, 0], [0, r, 0], [0, -r, 0], [0, 0, r], [0, 0, -r]],
                 np.float64) + c
    t = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    return v, np.asarray(t, np.int32)


def icosphere(center: Sequence[float], r: float, subdiv: int = 2) -> Mesh:
    """Subdivided icosahedron with its vertices on the sphere (convex)."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p),
         (0, 1, -p), (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    v = [np.asarray(a, np.float64) / np.linalg.norm(a) for a in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                w = v[a] + v[b]
                v.append(w / np.linalg.norm(w))
                mid[key] = len(v) - 1
            return mid[key]

        f = [g for (a, b, c) in f
             for g in ((a, m(a, b), m(a, c)), (b, m(b, c), m(a, b)), (c, m(a, c), m(b, c)),
                       (m(a, b), m(b, c), m(a, c)))]
    return np.asarray(v) * r + np.asarray(center, np.float64), np.asarray(f, np.int32)


def torus_mesh(center: Sequence[float], R: float, r: float, nu: int = 48, nv: int = 24) -> Mesh:
    """Closed torus, axis along y: nu x nv vertex grid, two triangles per
    quad (genus one, not convex)."""
    cx, cy, cz = (float(a) for a in center)
    th = 2.0 * np.pi * np.arange(nu) / nu
    ph = 2.0 * np.pi * np.arange(nv) / nv
    T, P = np.meshgrid(th, ph, indexing="ij")
    v = np.stack([cx + (R + r * np.cos(P)) * np.cos(T), cy + r * np.sin(P),
                  cz + (R + r * np.cos(P)) * np.sin(T)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    t = np.concatenate([np.stack([a, d, c], -1).reshape(-1, 3), np.stack([a, c, b], -1).reshape(-1, 3)])
    return v, t.astype(np.int32)


PROCEDURAL = {"box": box_mesh, "octahedron": octahedron, "icosphere": icosphere,
              "torus": torus_mesh}


def make_mesh(spec) -> Mesh:
    """A mesh from an OBJ path, a (vertices, triangles) pair, or a dict
    {"mesh": procedural name, **arguments}."""
    if isinstance(spec, (str, Path)):
        return read_obj(spec)
    if isinstance(spec, dict):
        kw = dict(spec)
        return PROCEDURAL[kw.pop("mesh")](**kw)
    v, t = spec
    return np.asarray(v, np.float64).reshape(-1, 3), np.asarray(t, np.int32).reshape(-1, 3)


# ---- containment on the GPU -------------------------------------------------
def inside_mask(mesh: Mesh, ax: np.ndarray, ay: np.ndarray, az: np.ndarray) -> np.ndarray:
    """Boolean [nx, ny, nz]: lattice point (ax[i], ay[j], az[k]) inside the
    closed mesh (wsph_mesh_inside; az ascending)."""
    import torch
    from . import _native as N
    if not torch.cuda.is_available():
        raise RuntimeError("mesh sampling runs on the GPU (torch.cuda); there is no CPU path")
    lib = N.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    v, t = make_mesh(mesh)
    ax, ay, az = (np.ascontiguousarray(a, np.float64) for a in (ax, ay, az))
    if az.size > 1 and np.any(np.diff(az) <= 0):
        raise ValueError("az must be ascending")
    dv = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    dt = torch.from_numpy(np.ascontiguousarray(t)).to(dev)
    dx, dy, dz = (torch.from_numpy(a).to(dev) for a in (ax, ay, az))
    out = torch.zeros(ax.size * ay.size * az.size, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    N.check("wsph_mesh_inside", lib.wsph_mesh_inside(
        dv.data_ptr(), v.shape[0], dt.data_ptr(), t.shape[0], dx.data_ptr(), ax.size,
        dy.data_ptr(), ay.size, dz.data_ptr(), az.size, out.data_ptr()))
    return out.cpu().numpy().astype(bool).reshape(ax.size, ay.size, az.size)


def lattice_axes(mesh: Mesh, scale: float, spacing: float):
    """The scaled vertices and the np.arange axes over their bounds
    (model.py:16-24)."""
    v, t = make_mesh(mesh)
    v = v * scale
    lo, hi = v.min(axis=0), v.max(axis=0)
    axes = [np.arange(lo[a], hi[a], spacing) for a in range(3)]
    return (v, t), axes


def sample_points(mesh, scale: float = 1.0, spacing: float = 0.05) -> np.ndarray:
    """``_sample_points`` (model.py:14-47): the lattice points inside the
    scaled mesh, float64 [m, 3], in meshgrid ("ij") order."""
    (v, t), axes = lattice_axes(mesh, float(scale), float(spacing))
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    if grid.shape[0] == 0:
        return grid
    return grid[inside_mask((v, t), *axes).reshape(-1)]


def load_model(mesh, scale: float = 1.0, pos=np.zeros(3), rot=np.eye(3),
               spacing: float = 0.05) -> np.ndarray:
    """``load_model`` (model.py:50-60): sampled points, rotated by
    ``points @ rot`` and moved by ``pos``."""
    points = sample_points(mesh, float(scale), float(spacing))
    points @= np.array(rot)
    return points + np.array(pos)
