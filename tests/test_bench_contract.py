"""bench.py keeps its output contract: one JSON line with the metric, value,
unit, GPU count, step counts, time per step, the roofline object of the
dominant kernel pair and the CPU baselines; more than one GPU is refused
with a message."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _last_json(out: str) -> dict:
    lines = [l for l in out.splitlines() if l.startswith("{")]
    assert lines, out[-2000:]
    return json.loads(lines[-1])


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in _last_json(r.stdout)["error"]


@pytest.mark.gpu
def test_bench_line_fields():
    print("[gpu] bench.py short run (contract fields)", flush=True)
    r = subprocess.run([sys.executable, "-u", str(ROOT / "bench.py"), "--steps", "8", "--warmup", "2",
                        "--compare", "previous", "--autotune", "off", "--cpu-seconds", "2",
                        "--no-cpu-faithful"],
                       capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    d = _last_json(r.stdout)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 8 and d["warmup"] == 2
    assert d["value"] > 0 and d["higher_is_better"] is True and d["scaling"] == "weak"
    assert d["config"]["workload"] == "dambreak_1m" and d["dtype"] == "f32"
    assert abs(d["ms_per_step"] - d["config"]["fluid_particles"] * d["steps"] / d["value"]
               * 1e3 / d["steps"]) < 0.05 * d["ms_per_step"] or d["nonconverged_steps"] > 0
    rl = d["roofline"]
    assert rl["bound"] == "hbm" and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert rl["achieved"] > 0 and abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-12
    assert rl["traffic"] is None or rl["traffic"] > 0
    assert rl["traffic"] is None or rl["traffic_source"].startswith("profiles/")
    lim = rl["limiter"]
    assert lim["source"].startswith("profiles/")
    assert 0 < lim["j1"]["valu_issue"] <= 1 and 0 < lim["j2"]["td_busy"] <= 1
    assert rl["launches_timed"] > 0
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["cores"] == 1 and cb["value"] > 0 and cb["sample"]
    assert d["plain_starts"]["previous"]["value"] > 0
