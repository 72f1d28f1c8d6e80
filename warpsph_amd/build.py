"""Build libwsph.so in-tree with hipcc for gfx950.

    python -m warpsph_amd.build

Every translation unit is compiled with -ffp-contract=off (no fused
multiply-add, so float32 results follow the reference's operation order) and
with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libwsph.so"
OBJ_DIR = PKG / "_build"
SOURCES = ["capi.hip", "sph.hip", "grid.hip", "mlp.hip", "learn.hip", "mesh.hip", "scene_gen.cpp"]
ARCH = os.environ.get("WSPH_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if c.startswith("/") and Path(c).exists():
            return c
    return "hipcc"


def _compile(src: str, extra=(), obj_dir=None) -> Path:
    obj = (obj_dir or OBJ_DIR) / (Path(src).stem + ".o")
    cmd = [_hipcc(), *FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, out: Path = OUT, defines=()) -> Path:
    """Compile every HIP source and link libwsph.so (always rebuilds).
    ``defines`` are extra -D flags (tuning constants of the kernels)."""
    extra = [f"-D{d}" for d in defines]
    obj_dir = OBJ_DIR / ("_".join(defines) or "default")
    obj_dir.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda src: _compile(src, extra, obj_dir), SOURCES))
    out = Path(out)
    tmp = out.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
