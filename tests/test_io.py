"""USDA recorder and usd2ply round trip (host only)."""
import tarfile

import numpy as np

from warpsph_amd.usd import UsdPointRecorder, read_positions


def test_usda_roundtrip_and_ply_archive(tmp_path):
    from usd2ply import usd_to_ply
    rng = np.random.default_rng(0)
    frames = {t: rng.normal(size=(50, 3)).astype(np.float32) for t in range(3)}
    r = UsdPointRecorder(str(tmp_path / "run.usd"), fps=60)
    for t, pts in frames.items():
        r.begin_frame(t / 60)
        r.render_points("fluid", pts, radius=0.01)
        r.end_frame()
    r.save()
    back = read_positions(str(tmp_path / "run.usd"))["fluid"]
    for t, pts in frames.items():
        assert np.allclose(back[t], pts, rtol=1e-6, atol=1e-7)
    path = usd_to_ply(str(tmp_path / "run.usd"))
    with tarfile.open(path) as tar:
        names = sorted(m.name for m in tar.getmembers())
        assert names == ["frame_0000.ply", "frame_0001.ply", "frame_0002.ply"]
        data = tar.extractfile("frame_0001.ply").read()
    head, body = data.split(b"end_header\n", 1)
    assert b"element vertex 50" in head
    rec = np.frombuffer(body, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("o", "u1")])
    assert np.allclose(rec["x"], frames[1][:, 0], rtol=1e-6, atol=1e-7)


def test_demo_shim_has_only_scoped_device():
    import importlib.util
    from pathlib import Path
    spec = importlib.util.spec_from_file_location(
        "warp_shim", Path(__file__).resolve().parents[1] / "scenes" / "warp.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    assert m.__all__ == ["ScopedDevice"]
