"""Train the pressure predictor on dam-break data and evaluate it.

    python -m warpsph_amd.train_predictor --out warpsph_amd/pressure_net.npy

1. (features, target) pairs from the batch workload's scenes, solved together
   in one batched solver (previous-step start, pair logging on; ``--scenes``
   of them, ``--steps`` steps each,
   ``--per-step`` random rows kept per step; weakly coupled particles,
   |a_ii / a_ref| < 0.05, are dropped since the solver never asks the network
   about them);
2. Adam on random mini-batches (``--train-steps``, ``--batch``, ``--lr``)
   through the library's BF16 matrix-core kernels;
3. Jacobi iterations per step, previous-step start vs learned start, on the
   evaluation scene (``--eval-config``, ``--eval-steps``).
Progress goes to stdout, one line per stage; the result is one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import config as cfgmod
from . import learn
from . import scene as scn
from .mlp import PressureNet


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--per-step", type=int, default=1024)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-config", default="dambreak_1m")
    ap.add_argument("--eval-steps", type=int, default=500)
    ap.add_argument("--out", default=None, help="save the trained parameters (.npy)")
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--sequential", action="store_true",
                    help="generate pairs scene by scene instead of in one batched solver")
    a = ap.parse_args(argv)
    import torch

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    cfgs = scn.batch_configs(cfgmod.batch_256())[:a.scenes]
    if a.sequential:
        X, T = learn.generate_pairs(cfgs, a.steps, a.per_step, seed=a.seed,
                                    log=log if a.scenes <= 16 else None)
    else:
        X, T = learn.generate_pairs_batched(cfgs, a.steps, a.per_step, seed=a.seed, log=log)
    keep = X[:, 1].abs() >= 0.05
    X, T = X[keep].contiguous(), T[keep].contiguous()
    t_gen = time.time() - t0
    log(f"pairs: {X.shape[0]} rows from {len(cfgs)} scenes x {a.steps} steps in {t_gen:.1f}s; "
        f"target mean {float(T.mean()):.4f} std {float(T.std()):.4f}")
    # hold out 5% of the rows
    g = torch.Generator(device=X.device)
    g.manual_seed(a.seed + 1)
    perm = torch.randperm(X.shape[0], device=X.device, generator=g)
    nv = max(1, X.shape[0] // 20)
    Xv, Tv = X[perm[:nv]].contiguous(), T[perm[:nv]].contiguous()
    Xt, Tt = X[perm[nv:]].contiguous(), T[perm[nv:]].contiguous()
    net = PressureNet(max_batch=a.batch, params=np.load(a.init) if a.init else None,
                      seed=a.seed)
    base = float(0.5 * (Tv ** 2).mean())
    log(f"held-out loss of the zero predictor (carried pressure): {base:.4e}")
    t1 = time.time()
    losses = learn.train(net, Xt, Tt, a.train_steps, a.batch, a.lr, seed=a.seed,
                         log_every=max(1, a.train_steps // 10), log=log)
    t_train = time.time() - t1
    val = learn.eval_loss(net, Xv, Tv)
    log(f"trained {a.train_steps} steps in {t_train:.1f}s; held-out loss {val:.4e} "
        f"({val / max(base, 1e-30):.3f} of the carried pressure's)")
    if a.out:
        net.save(a.out)
        log(f"saved {a.out}")
    res = dict(rows=int(X.shape[0]), gen_s=round(t_gen, 1), train_s=round(t_train, 1),
               loss_first=float(losses[:100].mean()), loss_last=float(losses[-100:].mean()),
               heldout_loss=val, heldout_loss_carry=base)
    if a.eval_steps > 0:
        ecfg = cfgmod.preset(a.eval_config)
        t2 = time.time()
        res.update(learn.compare_starts(ecfg, net, a.eval_steps))
        res["eval_s"] = round(time.time() - t2, 1)
        res["eval_config"] = a.eval_config
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
