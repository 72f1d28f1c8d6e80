"""USDA text recorder of particle positions.

Stands in for warp.render.UsdRenderer as the reference uses it
(IISPH.py:823-827, 1373-1406): one PointInstancer per named point set, with
positions sampled at time code round(sim_time * fps). Written as plain USDA
text, so no USD library is needed to produce it.
"""
from __future__ import annotations

from typing import Dict

import numpy as np


class UsdPointRecorder:
    def __init__(self, stage_path: str, fps: int = 60):
        self.stage_path = stage_path
        self.fps = fps
        self.time = 0
        self.frames: Dict[str, Dict[int, np.ndarray]] = {}
        self.radius: Dict[str, float] = {}

    def begin_frame(self, sim_time: float) -> None:
        self.time = int(round(sim_time * self.fps))

    def render_points(self, name: str, points, radius: float = 0.01) -> None:
        self.frames.setdefault(name, {})[self.time] = np.asarray(points, dtype=np.float32).copy()
        self.radius[name] = float(radius)

    def end_frame(self) -> None:
        pass

    def save(self) -> str:
        times = sorted({t for f in self.frames.values() for t in f})
        t0, t1 = (times[0], times[-1]) if times else (0, 0)
        out = ["#usda 1.0", "(", f"    startTimeCode = {t0}", f"    endTimeCode = {t1}",
               f"    timeCodesPerSecond = {self.fps}", '    upAxis = "Y"', ")", "",
               'def Xform "root"', "{"]
        for name, frames in self.frames.items():
            out.append(f'    def PointInstancer "{name}"')
            out.append("    {")
            n = max((f.shape[0] for f in frames.values()), default=0)
            out.append("        int[] protoIndices = [" + ", ".join(["0"] * n) + "]")
            out.append("        point3f[] positions.timeSamples = {")
            for t in sorted(frames):
                pts = ", ".join(f"({x:.7g}, {y:.7g}, {z:.7g})" for x, y, z in frames[t])
                out.append(f"            {t}: [{pts}],")
            out.append("        }")
            out.append(f"        float radius = {self.radius[name]:.7g}")
            out.append("    }")
        out.append("}")
        with open(self.stage_path, "w") as fh:
            fh.write("\n".join(out) + "\n")
        return self.stage_path


def read_positions(stage_path: str) -> Dict[str, Dict[int, np.ndarray]]:
    """Parse the point sets written by :class:`UsdPointRecorder`."""
    import re
    res: Dict[str, Dict[int, np.ndarray]] = {}
    cur = None
    num = r"[-+0-9.eE]+|nan|inf|-inf"
    trip = re.compile(rf"\(({num}), ({num}), ({num})\)")
    with open(stage_path) as fh:
        for line in fh:
            m = re.match(r'\s*def PointInstancer "(\w+)"', line)
            if m:
                cur = m.group(1)
                res[cur] = {}
                continue
            m = re.match(r"\s*(-?\d+): \[(.*)\],\s*$", line)
            if m and cur is not None:
                vals = trip.findall(m.group(2))
                res[cur][int(m.group(1))] = np.array(vals, dtype=np.float32).reshape(-1, 3)
    return res
