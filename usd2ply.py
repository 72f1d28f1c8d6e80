"""USD point recordings -> per-frame PLY files in one archive.

Same entry point as the reference's usd2ply.py (``usd_to_ply(usd_file_path)``,
usd2ply.py:45-92): every time code of the first point set becomes
``frame_NNNN.ply`` (vertex x, y, z float32 and obj_id uint8) inside
``<name>.tar.zst``. The USD library, zstandard and joblib are optional here:
without ``pxr`` the USDA text written by warpsph_amd.usd is parsed directly,
and without ``zstandard`` the archive is an uncompressed ``<name>.tar``.
"""
from __future__ import annotations

import io
import os
import tarfile

import numpy as np


def _ply_bytes(points: np.ndarray) -> bytes:
    pts = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    rec = np.zeros(pts.shape[0], dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                                        ("obj_id", "u1")])
    rec["x"], rec["y"], rec["z"] = pts[:, 0], pts[:, 1], pts[:, 2]
    head = ("ply\nformat binary_little_endian 1.0\n"
            f"element vertex {pts.shape[0]}\n"
            "property float x\nproperty float y\nproperty float z\n"
            "property uchar obj_id\nend_header\n").encode()
    return head + rec.tobytes()


def _read_frames(usd_file_path: str):
    """{time code: positions} of the first point set in the stage."""
    try:
        from pxr import Usd, UsdGeom  # type: ignore
    except ImportError:
        from warpsph_amd.usd import read_positions
        sets = read_positions(usd_file_path)
        return next(iter(sets.values())) if sets else {}
    stage = Usd.Stage.Open(usd_file_path)
    frames = {}
    for prim in stage.Traverse():
        if prim.IsA(UsdGeom.PointInstancer):
            attr = UsdGeom.PointInstancer(prim).GetPositionsAttr()
            t0, t1 = int(stage.GetStartTimeCode()), int(stage.GetEndTimeCode())
            for t in range(t0, t1 + 1):
                pos = attr.Get(t)
                if pos:
                    frames[t] = np.array(pos, dtype=np.float32)
            break
    return frames


def usd_to_ply(usd_file_path: str) -> str:
    frames = _read_frames(usd_file_path)
    base = os.path.splitext(os.path.basename(usd_file_path))[0]
    out_dir = os.path.dirname(usd_file_path)
    buf = io.BytesIO()
    with tarfile.open(fileobj=buf, mode="w") as tar:
        t0 = min(frames) if frames else 0
        for t in sorted(frames):
            data = _ply_bytes(frames[t])
            info = tarfile.TarInfo(f"frame_{t - t0:04d}.ply")
            info.size = len(data)
            tar.addfile(info, io.BytesIO(data))
    try:
        import zstandard as zstd  # type: ignore
        path = os.path.join(out_dir, f"{base}.tar.zst")
        with open(path, "wb") as fh:
            fh.write(zstd.ZstdCompressor(level=3).compress(buf.getvalue()))
    except ImportError:
        path = os.path.join(out_dir, f"{base}.tar")
        with open(path, "wb") as fh:
            fh.write(buf.getvalue())
    print(f"Export completed: {path} ({len(frames)} frames)")
    return path


if __name__ == "__main__":
    import sys
    usd_to_ply(sys.argv[1])
