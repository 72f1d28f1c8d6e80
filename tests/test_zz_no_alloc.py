"""wsph_mesh_inside_noalloc against wsph_mesh_inside, through the C-ABI.

The variant that allocates nothing (its scratch is one fixed device array per
GPU, kept between calls) must give the allocating entry point's result bit
for bit, also when a call follows another on the same array with a smaller
input (whatever the last call left there must not matter), and must refuse
work beyond its scratch with WSPH_E_CAPACITY and still serve the next call.
"""
import sys

import numpy as np
import pytest

from warpsph_amd import _native as N
from warpsph_amd import mesh as M

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _axes(lo, hi, d):
    return [np.arange(lo[a], hi[a], d) for a in range(3)]


def _classify(fn, mesh, axes):
    """(status, inside[nx, ny, nz]) of one call of entry point ``fn``."""
    import torch
    lib = N.load()
    v, t = M.make_mesh(mesh)
    ax, ay, az = (np.ascontiguousarray(a, np.float64) for a in axes)
    dv = torch.from_numpy(np.ascontiguousarray(v)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(t)).cuda()
    dx, dy, dz = (torch.from_numpy(a).cuda() for a in (ax, ay, az))
    out = torch.full((ax.size * ay.size * az.size,), 7, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(lib, fn)(dv.data_ptr(), v.shape[0], dt.data_ptr(), t.shape[0], dx.data_ptr(),
                          ax.size, dy.data_ptr(), ay.size, dz.data_ptr(), az.size, out.data_ptr())
    torch.cuda.synchronize()
    return rc, out.cpu().numpy().reshape(ax.size, ay.size, az.size)


LARGE = (M.torus_mesh((0, 0, 0), 0.3, 0.1, 96, 24),
         _axes((-0.42, -0.12, -0.42), (0.42, 0.12, 0.42), 0.004))
SMALL = (M.icosphere((0.5, 0.4, 0.45), 0.3, 2), _axes((0.1, 0, 0.05), (0.9, 0.8, 0.85), 0.021))


def test_noalloc_equals_mesh_inside_on_consecutive_calls():
    for name, (mesh, axes) in (("large torus", LARGE), ("small icosphere", SMALL),
                               ("large torus again", LARGE), ("small icosphere again", SMALL)):
        _announce(f"no-alloc mesh test: {name}, {[a.size for a in axes]} lattice")
        rc, want = _classify("wsph_mesh_inside", mesh, axes)
        assert rc == N.WSPH_OK
        rc, got = _classify("wsph_mesh_inside_noalloc", mesh, axes)
        assert rc == N.WSPH_OK, N.load().wsph_last_error()
        assert want.any() and not want.all()
        assert np.array_equal(got, want), f"{name}: {int((got != want).sum())} points differ"


def test_noalloc_capacity_error_then_next_call():
    """Twelve stacked boxes over a 1024 x 1024 column lattice cross every
    column 24 times: 25 M crossings (8 bytes each) exceed the 64 MiB array,
    which the count pass finds; the call fails cleanly and the next, smaller
    call is still exact."""
    boxes = [M.box_mesh((0.0, 0.0, 0.1 * k), (1.0, 1.0, 0.1 * k + 0.05)) for k in range(12)]
    v = np.concatenate([b[0] for b in boxes])
    t = np.concatenate([b[1] + 8 * k for k, b in enumerate(boxes)])
    axes = [np.linspace(0.0005, 0.9995, 1024), np.linspace(0.0005, 0.9995, 1024),
            np.array([0.02, 0.5, 0.93, 1.2])]
    _announce("no-alloc mesh test: crossings beyond the scratch array")
    rc, _ = _classify("wsph_mesh_inside_noalloc", (v, t), axes)
    assert rc == N.WSPH_E_CAPACITY
    assert b"column crossings" in N.load().wsph_last_error()
    rc, want = _classify("wsph_mesh_inside", *SMALL)
    rc2, got = _classify("wsph_mesh_inside_noalloc", *SMALL)
    assert rc == rc2 == N.WSPH_OK
    assert np.array_equal(got, want)
