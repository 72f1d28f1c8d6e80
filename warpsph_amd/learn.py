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

from . import _native as N
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


def generate_pairs_batched(cfgs, steps: int, per_step: int, seed: int = 0,
                           device: str = "cuda:0", log=None, shape: Optional[dict] = None):
    """As ``generate_pairs``, with all scenes in one batched solver
    (``BatchIISPH``): per step, ``per_step`` random rows of every scene."""
    from .batch import BatchIISPH
    torch = _torch()
    cfgs = [c.replace(warm_start=cfgmod.WARM_PREVIOUS, **(shape or {})) for c in cfgs]
    b = BatchIISPH(cfgs, device=device)
    N.check("wsph_set_pair_logging", b._lib.wsph_set_pair_logging(b._h, 1))
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    counts = torch.as_tensor(b.counts, dtype=torch.float64, device=device)
    offs = torch.as_tensor(b.offsets[:-1], dtype=torch.int64, device=device)
    take = torch.minimum(torch.full_like(counts, per_step), counts).to(torch.int64)
    f = torch.empty((b.n, 8), dtype=torch.float32, device=device)
    t = torch.empty(b.n, dtype=torch.float32, device=device)
    feats, targs = [], []
    t0 = time.time()
    for k in range(steps):
        b.step()
        torch.cuda.synchronize()
        N.check("wsph_get_pairs", b._lib.wsph_get_pairs(b._h, f.data_ptr(), t.data_ptr()))
        u = torch.rand((b.nscenes, per_step), device=device, generator=gen, dtype=torch.float64)
        idx = (offs[:, None] + (u * counts[:, None]).to(torch.int64))
        # scenes smaller than per_step contribute each row once
        keep = torch.arange(per_step, device=device)[None, :] < take[:, None]
        small = counts < per_step
        if bool(small.any()):
            idx = torch.where(small[:, None], offs[:, None] + torch.arange(per_step, device=device)[None, :],
                              idx)
        idx = idx[keep]
        feats.append(f[idx])
        targs.append(t[idx])
        if log and (k + 1) % 50 == 0:
            c = b.counters()
            log(f"batch pairs: step {k + 1}/{steps}, iters/step {c['iterations'] / max(c['steps'], 1):.2f}, "
                f"{time.time() - t0:.1f}s")
    torch.cuda.synchronize()
    b.close()
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
