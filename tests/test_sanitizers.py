"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the CPU code.

The C oracle (oracle/iisph_oracle.c, oracle/mesh_oracle.c) and the host
scene generator (warpsph_amd/csrc/scene_gen.cpp) are compiled with gcc/g++
and -fsanitize=address,undefined (host code only: no GPU code is built
here) together with small drivers under tests/sanitize/ that exercise every
entry point, and run as child processes; any sanitizer report fails the
test (-fno-sanitize-recover makes the first one fatal).
"""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def _run(exe: Path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, f"{exe.name} failed ({r.returncode}):\n{r.stdout}\n{r.stderr[-4000:]}"
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_san"
    subprocess.run(["gcc", "-std=c11", "-ffp-contract=off", "-fopenmp", *SAN, "-o", str(exe),
                    str(ROOT / "tests/sanitize/oracle_driver.c"),
                    str(ROOT / "oracle/iisph_oracle.c"), str(ROOT / "oracle/mesh_oracle.c"),
                    "-lm"], check=True, capture_output=True)
    assert "oracle driver ok" in _run(exe)


def test_scene_gen_under_sanitizers(tmp_path):
    exe = tmp_path / "scene_gen_san"
    subprocess.run(["g++", "-std=c++17", *SAN, "-o", str(exe),
                    str(ROOT / "tests/sanitize/scene_gen_driver.cpp"),
                    str(ROOT / "warpsph_amd/csrc/scene_gen.cpp")], check=True, capture_output=True)
    assert "scene_gen driver ok" in _run(exe)
