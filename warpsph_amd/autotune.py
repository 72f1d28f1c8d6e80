"""Launch-shape autotuner for the IISPH step.

Times a few steps of the actual workload for each candidate shape: threads per
block of the neighbour passes, Jacobi driver (device-driven chunks of 1, 2 or
4 iterations, the same with the step's front half replayed as a captured
hipGraph holding 2 or 4 iterations, or the persistent cooperative kernel at 1
or 2 blocks per CU) and cell layout (Morton or linear). Every candidate computes the same numbers
(the neighbour order and the float operations do not depend on the shape),
so the candidates run identical work from the same initial state and only
their wall time differs; the fastest few are timed again before one is
chosen. Results are cached per device and workload in a
JSON file (default ``~/.cache/warpsph_amd/autotune.json``).
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Callable, Dict, Iterable, List, Optional

from . import config as cfgmod

DEFAULT_CACHE = Path(os.environ.get("WSPH_AUTOTUNE_CACHE",
                                    Path.home() / ".cache" / "warpsph_amd" / "autotune.json"))

SHAPE_KEYS = ("block", "jacobi_mode", "chunk_iters", "coop_blocks_per_cu", "cell_layout",
              "graph_iters")


def candidates(blocks: Iterable[int] = (128, 256, 512),
               layouts: Iterable[int] = (cfgmod.LAYOUT_MORTON, cfgmod.LAYOUT_LINEAR)) -> List[dict]:
    drivers = [dict(jacobi_mode=cfgmod.JACOBI_DEVICE, chunk_iters=c, coop_blocks_per_cu=0,
                    graph_iters=0) for c in (1, 2, 4)]
    drivers += [dict(jacobi_mode=cfgmod.JACOBI_DEVICE, chunk_iters=2, coop_blocks_per_cu=0,
                     graph_iters=g) for g in (2, 4)]
    drivers += [dict(jacobi_mode=cfgmod.JACOBI_PERSISTENT, chunk_iters=0, coop_blocks_per_cu=c,
                     graph_iters=0) for c in (1, 2)]
    return [dict(block=b, cell_layout=l, **d) for b in blocks for l in layouts for d in drivers]


def _device_name() -> str:
    import torch
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _load(cache: Path) -> dict:
    try:
        return json.loads(cache.read_text())
    except (OSError, ValueError):
        return {}


def _time_shape(make, cfg: cfgmod.SolverConfig, shape: dict, warmup: int, steps: int) -> float:
    """Milliseconds per step of ``steps`` steps after ``warmup``, on a fresh solver."""
    import torch
    s = make(cfg.replace(**shape))
    try:
        for _ in range(warmup):
            s.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3
    finally:
        s.close()


def tune(make: Callable[[cfgmod.SolverConfig], object], cfg: cfgmod.SolverConfig, key: str,
         cands: Optional[List[dict]] = None, steps: int = 10, warmup: int = 3,
         cache: Optional[Path] = DEFAULT_CACHE, force: bool = False, log=None,
         finalists: int = 4, rounds: int = 2) -> Dict:
    """Best shape for ``make(cfg.replace(**shape))`` by wall time of ``steps``
    steps after ``warmup``; returns {"shape": ..., "ms_per_step": ..., "table": [...],
    "final": [...]}. ``make`` builds a solver with ``step()`` and ``close()``.

    Every candidate is timed once, in order; the ``finalists`` fastest are then
    timed ``rounds`` more times each, in alternating order, and the one with
    the lowest of those timings wins. One timing per shape can differ from the
    next by a few percent, more than the gap between the leading shapes."""
    full_key = f"{_device_name()}|{key}"
    if cache is not None and not force:
        hit = _load(Path(cache)).get(full_key)
        if hit:
            return dict(hit, cached=True)
    table = []
    for shape in cands if cands is not None else candidates():
        ms = _time_shape(make, cfg, shape, warmup, steps)
        table.append(dict(shape=shape, ms_per_step=ms))
        if log:
            log(f"autotune {key}: {shape} {ms:.3f} ms/step")
    top = [r["shape"] for r in sorted(table, key=lambda r: r["ms_per_step"])[:max(finalists, 1)]]
    final = [dict(shape=sh, ms=[]) for sh in top]
    if finalists > 0 and rounds > 0 and len(top) > 1:
        for rnd in range(rounds):
            for f in (final if rnd % 2 == 0 else final[::-1]):
                f["ms"].append(_time_shape(make, cfg, f["shape"], warmup, steps))
                if log:
                    log(f"autotune {key} finalist: {f['shape']} {f['ms'][-1]:.3f} ms/step")
        best = min(final, key=lambda f: min(f["ms"]))
        best_shape, best_ms = best["shape"], min(best["ms"])
    else:
        best_shape = top[0]
        best_ms = min(r["ms_per_step"] for r in table)
    out = dict(shape=best_shape, ms_per_step=best_ms, table=table, final=final, steps=steps,
               warmup=warmup)
    if cache is not None:
        cache = Path(cache)
        data = _load(cache)
        data[full_key] = out
        cache.parent.mkdir(parents=True, exist_ok=True)
        cache.write_text(json.dumps(data, indent=1))
    return dict(out, cached=False)


def apply(cfg: cfgmod.SolverConfig, shape: dict) -> cfgmod.SolverConfig:
    return cfg.replace(**{k: shape[k] for k in SHAPE_KEYS if k in shape})
