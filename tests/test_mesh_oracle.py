"""The oracle of the mesh containment test (model.py:14-47's
Trimesh.contains, restated as ray parity along +z with the top-left rule),
pinned against methods that share nothing with it: the analytic interior of
a box and of an octahedron, half-space tests for a convex polyhedron, and the
generalized winding number (a sum of solid angles) for a torus. CPU only."""
import numpy as np
import pytest

from warpsph_amd import mesh as M


def _axes(lo, hi, d):
    return [np.arange(lo[a], hi[a], d) for a in range(3)]


def _grid(axes):
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1)


def test_box_unaligned_lattice(orc):
    v, t = M.box_mesh((0.1, 0.2, 0.3), (0.73, 0.61, 0.95))
    axes = _axes((-0.013, 0.011, 0.017), (0.9, 0.8, 1.1), 0.05)
    got = orc.mesh_inside(v, t, *axes)
    g = _grid(axes)
    want = np.all((g > v.min(0)) & (g < v.max(0)), axis=-1)
    assert got.any() and np.array_equal(got, want)


def test_octahedron_rays_through_edges_and_vertices(orc):
    """Lattice columns pass exactly through the projected edges (x = 0,
    y = 0, |x| + |y| = r) and through the top and bottom vertices, so every
    tie of the top-left rule occurs; points off the surface must still be
    classified exactly."""
    r = 1.0
    v, t = M.octahedron((0.0, 0.0, 0.0), r)
    axes = [np.arange(-1.5, 1.5001, 0.25)] * 3
    got = orc.mesh_inside(v, t, *axes)
    g = _grid(axes)
    l1 = np.abs(g).sum(-1)
    assert np.array_equal(got[l1 < r], np.ones(int((l1 < r).sum()), bool))
    assert not got[l1 > r].any()


def test_icosphere_half_spaces(orc):
    v, t = M.icosphere((0.5, 0.4, 0.45), 0.3, subdiv=2)
    axes = _axes((0.1, 0.0, 0.05), (0.9, 0.8, 0.85), 0.021)
    got = orc.mesh_inside(v, t, *axes)
    g = _grid(axes).reshape(-1, 3)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    n *= np.sign(((a - v.mean(0)) * n).sum(1))[:, None]  # outward
    s = g @ n.T - (a * n).sum(1)[None, :]  # signed distance to each face plane
    clear = np.abs(s).min(1) > 1e-9
    want = np.all(s < 0, axis=1)
    assert np.array_equal(got.reshape(-1)[clear], want[clear])
    assert want[clear].sum() > 1000


def _winding_number(v, t, pts):
    """Sum of the triangles' solid angles / 4 pi (Van Oosterom and
    Strackee's formula)."""
    w = np.zeros(pts.shape[0])
    for tri in t:
        a, b, c = (v[k][None, :] - pts for k in tri)
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
        det = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = (la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", b, c) * la
               + np.einsum("ij,ij->i", c, a) * lb)
        w += 2.0 * np.arctan2(det, den)
    return w / (4.0 * np.pi)


def test_torus_winding_number(orc):
    v, t = M.torus_mesh((0.0, 0.0, 0.0), 0.3, 0.1, nu=24, nv=12)
    axes = _axes((-0.42, -0.13, -0.41), (0.42, 0.13, 0.42), 0.0299)
    got = orc.mesh_inside(v, t, *axes)
    g = _grid(axes).reshape(-1, 3)
    w = np.abs(_winding_number(v, t, g))
    clear = np.minimum(np.abs(w), np.abs(w - 1.0)) < 1e-6
    assert clear.mean() > 0.99
    want = w > 0.5
    assert np.array_equal(got.reshape(-1)[clear], want[clear])
    assert want[clear].sum() > 100 and (~want[clear]).sum() > 100
    # the hole: the axis of the torus is outside
    i0 = np.argmin(np.abs(axes[0]))
    k0 = np.argmin(np.abs(axes[2]))
    assert not got[i0, :, k0].any()


def test_bad_index_is_an_error(orc):
    with pytest.raises(ValueError):
        orc.mesh_inside(np.zeros((3, 3)), np.array([[0, 1, 3]]), [0.0], [0.0], [0.0])


def test_read_obj(tmp_path):
    """Quads, slashed index forms and negative indices parse into the same
    closed box as box_mesh."""
    v, _ = M.box_mesh((0, 0, 0), (1, 2, 3))
    lines = [f"v {a} {b} {c}" for a, b, c in v]
    lines += ["vn 0 0 1", "f 1//1 3//1 4//1 2//1", "f 5/1/1 6/1/1 8/1/1 7/1/1",
              "f 1 2 6 5", "f -6 -2 -1 -5", "f 1 5 7 3", "f 2 4 8 6"]
    p = tmp_path / "box.obj"
    p.write_text("\n".join(lines) + "\n")
    vv, tt = M.read_obj(p)
    assert np.array_equal(vv, v) and tt.shape == (12, 3)
    vb, tb = M.box_mesh((0, 0, 0), (1, 2, 3))
    key = lambda tris: sorted(tuple(sorted(x)) for x in tris.tolist())
    assert key(tt) == key(tb)
