"""Training driver of the pressure predictor.

1. ``generate_pairs``: run dam-break scenes with the previous-step start and
   pair logging on; after every step keep a random sample of the particles'
   (features, target) rows, target = (converged p - carried p) / P_ref.
2. ``train``: Adam on random mini-batches of those rows, through the
   library's own BF16 matrix-core forward and backward kernels.
3. ``compare_starts``: run one scene twice, previous-step start and learned
   start, and report Jacobi iterations per step and the start-value errors.

torch holds the device tensors and draws the random indices; every network
evaluation and gradient runs in libwsph.so.
"""
from __future__ import annotations

import time
from typing import Iterable, Optional

import numpy as np

from . import config as cfgmod
from . import scene as scn
from .mlp import N_IN, PressureNet


def _torch():
    import torch
    return torch


def generate_pairs(cfgs: Iterable[cfgmod.SolverConfig], steps: int, per_step: int,
                   seed: int = 0, device: str = "cuda:0", log=None):
    """Concatenated (features [N, 8], targets [N]) device tensors from
    ``steps`` steps of each configuration, ``per_step`` rows sampled per step
    (all rows when a scene has fewer)."""
    from .solver import IISPH
    torch = _torch()
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    feats, targs = [], []
    for k, cfg in enumerate(cfgs):
        cfg = cfg.replace(warm_start=cfgmod.WARM_PREVIOUS)
        s = IISPH(stage_path=None, config=cfg, device=device)
        s.set_pair_logging(True)
        t0 = time.time()
        for _ in range(steps):
            s.step()
            f, t = s.pairs()
            if s.n > per_step:
                idx = torch.randint(0, s.n, (per_step,), device=device, generator=gen)
                f, t = f[idx], t[idx]
            feats.append(f)
            targs.append(t)
        torch.cuda.synchronize()
        if log:
            c = s.counters()
            log(f"scene {k}: n={s.n} steps={steps} iters/step={c['iterations'] / max(c['steps'], 1):.2f} "
                f"{time.time() - t0:.1f}s")
        s.close()
    return torch.cat(feats), torch.cat(targs)


def train(net: PressureNet, X, T, steps: int, batch: int = 65536, lr: float = 1e-3,
          seed: int = 0, log_every: int = 0, log=None):
    """Adam on uniformly drawn mini-batches; returns the loss of every step."""
    torch = _torch()
    if X.shape[0] == 0:
        raise ValueError("no training rows")
    batch = min(batch, net.max_batch)
    gen = torch.Generator(device=X.device)
    gen.manual_seed(seed)
    losses = np.zeros(steps, dtype=np.float64)
    for k in range(steps):
        idx = torch.randint(0, X.shape[0], (batch,), device=X.device, generator=gen)
        losses[k] = net.train_step(X[idx], T[idx], lr=lr)
        if log and log_every and (k + 1) % log_every == 0:
            log(f"train step {k + 1}: loss {losses[max(0, k - log_every + 1):k + 1].mean():.4e}")
    return losses


def eval_loss(net: PressureNet, X, T, batch: int = 65536) -> float:
    """Mean 0.5 (net - t)^2 over all rows (forward only)."""
    torch = _torch()
    tot, n = 0.0, 0
    for a in range(0, X.shape[0], batch):
        y = net.forward(X[a:a + batch])
        tot += 0.5 * float(((y - T[a:a + batch]) ** 2).sum())
        n += y.shape[0]
    return tot / max(n, 1)


def run_scene(cfg: cfgmod.SolverConfig, steps: int, net: Optional[PressureNet] = None,
              device: str = "cuda:0") -> dict:
    """Run ``steps`` steps; return the counters and the per-step history."""
    from .solver import IISPH
    s = IISPH(stage_path=None, config=cfg, device=device, predictor=net)
    for _ in range(steps):
        s.step()
    out = dict(counters=s.counters(), history=s.history(steps))
    s.close()
    return out


def compare_starts(cfg: cfgmod.SolverConfig, net: PressureNet, steps: int,
                   device: str = "cuda:0") -> dict:
    """Iterations per step and start-value errors, previous-step start vs
    learned start, on the same scene."""
    prev = run_scene(cfg.replace(warm_start=cfgmod.WARM_PREVIOUS), steps, None, device)
    lrn = run_scene(cfg.replace(warm_start=cfgmod.WARM_LEARNED), steps, net, device)
    h = lrn["history"]
    used = sum(1 for r in h if r["start_source"] == cfgmod.WARM_LEARNED)

    def per_step(r):
        c = r["counters"]
        return c["iterations"] / max(c["steps"], 1)

    # the physics must not move: last-step statistics of both runs
    hp = prev["history"][-1] if prev["history"] else {}
    hl = h[-1] if h else {}
    drift = {k: (hl[k] - hp[k]) / hp[k] if hp.get(k) else 0.0
             for k in ("kinetic_energy", "mean_height", "surface_height")}
    return dict(
        final_rel_diff=drift,
        iters_previous=per_step(prev), iters_learned=per_step(lrn),
        nonconverged_previous=prev["counters"]["nonconverged"],
        nonconverged_learned=lrn["counters"]["nonconverged"],
        err_learned=float(np.mean([r["start_err_learned"] for r in h])) if h else 0.0,
        err_carry=float(np.mean([r["start_err_carry"] for r in h])) if h else 0.0,
        steps_on_network=used, steps=steps)
