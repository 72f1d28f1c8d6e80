"""Throughput of the IISPH step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--preset dambreak_1m]

Workload: the 1,000,000-particle dam break (100^3 fluid lattice at 0.01 m,
+-1% jitter from seed 0, 2x1x1 m box with one wall layer, dt = 1e-3,
XSPH 0.01, relaxed Jacobi to a mean density error <= 1e-3). W untimed steps
run first; then exactly K steps are timed between two device
synchronisations with the particle state resident in HBM. ``value`` counts
fluid particle-steps per second over the steps whose pressure solve
converged (residual <= eta).

The headline run uses the learned pressure start (the predictor network in
warpsph_amd/pressure_net.npy, trained by warpsph_amd.train_predictor); the
same K steps are then run from fresh state with the plain starts, the
previous step's pressure and p = 0 (the reference's), and their throughput
and Jacobi iterations per step are reported beside it.

The roofline figure covers the dominant kernel pair, J1 + J2 of the Jacobi
loop: its algorithmic bytes are 92 B per fluid particle per A p evaluation
(J1: x, p, rho, d_ii read 32 B, s and A p partial written 16 B; J2: x, s,
A p, a_ii, p, rho_adv read 40 B, p_next written 4 B), and its duration is
measured with HIP events on the solver's stream around the J1 and J2
launches of every timed step's first Jacobi iteration.

The CPU baseline times the float32 CPU oracle (one thread) on a bounded
sample: the GPU's particle state at the end of the timed run, advanced a
few steps. Beside it: the same on every host core, with the reference's hash
dims on every core, and one step with the reference's hash dims on one thread
(the reference's algorithm as Warp's serial CPU launches run it;
--no-cpu-faithful skips that one).

``--preset mesh_sample`` measures the model sampler instead (model.py:14-47):
the lattice points inside a closed mesh, a 36,864-triangle torus
(R = 0.3, r = 0.1) on its 400 x 100 x 400 lattice at 0.002 m, classified by
wsph_mesh_inside with the mesh and lattice resident on the device; a "step"
is one whole classification. Its roofline covers the longest of the five
kernels (HIP event durations from the library), priced by the bytes it must
move: each input read once, each output written once.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
WARM = {}
BYTES_PER_AP_EVAL = 92  # per fluid particle, J1 48 + J2 44
BYTES_PER_STEP_FIXED = 264  # grid 84 + density 16 + v_adv 40 + assembly 48 + integrate 76
MESH_SAMPLE = dict(mesh="torus", center=(0.0, 0.0, 0.0), R=0.3, r=0.1, nu=192, nv=96)
MESH_SPACING = 0.002


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--preset", default="dambreak_1m")
    ap.add_argument("--jacobi", choices=["persistent", "host", "device"], default="device")
    ap.add_argument("--chunk", type=int, default=0, help="device-loop iterations per chunk")
    ap.add_argument("--warm-start", choices=["zero", "previous", "learned"], default="learned")
    ap.add_argument("--net", default=str(ROOT / "warpsph_amd" / "pressure_net.npy"),
                    help="predictor parameters for the learned start")
    ap.add_argument("--compare", default="previous,zero",
                    help="plain starts to run after the headline run (comma list, '' for none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-faithful", action="store_true",
                    help="skip the one-step, one-thread CPU timing with the reference's hash "
                         "dims (~36 s at 1M particles)")
    ap.add_argument("--layout", type=int, default=-1, help="cell layout: 0 linear, 1 Morton")
    ap.add_argument("--block", type=int, default=0, help="threads per block (0: solver default)")
    ap.add_argument("--graph-iters", type=int, default=-1,
                    help="Jacobi iterations in the captured step graph (0: no graph; -1: default)")
    ap.add_argument("--coop-per-cu", type=int, default=0,
                    help="persistent kernel blocks per CU (0: solver default)")
    ap.add_argument("--trace-steps", type=int, default=0,
                    help="print wall time and solver statistics of this many single steps first")
    ap.add_argument("--autotune", choices=["off", "cached", "force"], default="cached",
                    help="launch shapes from the autotuner (its cache, or a fresh sweep)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def calibrate_copy(mib: int) -> None:
    """A plain device-to-device copy of known size, to scale the FETCH_SIZE /
    WRITE_SIZE counters of a profiling run."""
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 4, "launches": 5}), flush=True)


def traffic_from_profile(workload: str):
    """(HBM bytes per launch of the Jacobi kernel pair, the summary they come
    from) out of the committed rocprofv3 counter summary, if one exists for
    this workload."""
    f = ROOT / "profiles" / f"jacobi_{workload}_traffic.json"
    if not f.exists():
        return None, None
    try:
        d = json.loads(f.read_text())
        return d.get("bytes_per_launch"), d.get("source")
    except Exception:
        return None, None


def limiter_from_profile(workload: str):
    """What bounds J1 and J2, from the committed counter summary of the
    vector-memory pipeline (profsum pipeline): VALU issue fraction and TA /
    TD busy fractions per kernel, if a summary exists for this workload."""
    f = ROOT / "profiles" / f"jacobi_{workload}_pipeline.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
    except Exception:
        return None
    out = {"source": d.get("source"), "note": d.get("note")}
    for k, short in (("k_j1_dev", "j1"), ("k_j2_dev", "j2")):
        if k in d:
            out[short] = {q: round(d[k][q], 3) for q in ("valu_issue", "ta_busy", "td_busy",
                                                        "l1_hit", "l2_hit") if q in d[k]}
    return out


def _nproc() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(cfg, x, v, budget_s: float, threads: int = 1, boundary=None):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from warpsph_amd import scene as scn
    O.set_threads(threads)
    sc = scn.build(cfg)
    sc.x = x
    O_s = O.OracleSolver(cfg, sc)
    O_s.v = v.copy()
    if boundary is not None:  # moved model: its positions, velocities and the clock
        O_s.xb, O_s.vb, O_s.sim_time = boundary[0].copy(), boundary[1].copy(), boundary[2]
    t0 = time.perf_counter()
    steps = 0
    while True:
        O_s.step(want_fields=False)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s * 0.5 or el * (steps + 1) / steps > budget_s:
            break
    el = time.perf_counter() - t0
    dims = scn.grid_dims_for(cfg, sc)[0]
    grid = (f"the reference's hash dims {dims[0]}x{dims[1]}x{dims[2]} (int(BOX/h/4))"
            if cfg.grid == "reference" else "domain-covering hash dims")
    return {"value": cfg_n(x) * steps / el, "unit": "fluid particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{cfg_n(x)}-particle state at the end of the GPU run, {steps} step(s), "
                      f"{el:.1f} s, {threads} thread(s) (OpenMP), float32 oracle restating the "
                      f"reference step with {grid}"}


def cfg_n(x):
    return int(x.shape[0])


PHASE_STEPS = 40


def phase_split(lib, h, steps: int = PHASE_STEPS) -> dict:
    """ms per step of the four phases, from ``steps`` further steps with phase
    events (timing mode 1). The timed run itself records only the J1/J2
    events of every 4th step (mode 2): an event between two kernels costs a
    few microseconds of device time, which phase events add to every step."""
    from warpsph_amd import _native as N
    N.check("wsph_set_timing", lib.wsph_set_timing(h, 1))
    for _ in range(steps):
        N.check("wsph_step", lib.wsph_step(h, None))
    tm = (C.c_double * 12)()
    N.check("wsph_last_timings", lib.wsph_last_timings(h, tm, 12))
    N.check("wsph_set_timing", lib.wsph_set_timing(h, 0))
    k = max(1, int(tm[6]))
    return {"neighbor_search": tm[0] / k, "predict_advection": tm[1] / k,
            "pressure_solve": tm[2] / k, "integrate": tm[3] / k, "steps": int(tm[6])}


def timed_run(cfg, net, args, trace: bool) -> dict:
    """W untimed steps, then K steps timed between device synchronisations."""
    import torch
    from warpsph_amd import _native as N
    from warpsph_amd.solver import IISPH
    t_setup = time.perf_counter()
    s = IISPH(stage_path=None, config=cfg, predictor=net)
    lib, h = s._lib, s._h
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s n={s.n} nb={s.boundary_n} "
          f"start={cfg.warm_start}", file=sys.stderr, flush=True)
    for k in range(args.trace_steps if trace else 0):
        st = N.WsphStepStats()
        t = time.perf_counter()
        rc = lib.wsph_step(h, C.byref(st))
        print(f"[bench] trace step {k}: rc={rc} {1e3 * (time.perf_counter() - t):.2f} ms "
              f"iters={st.iterations} residual={st.residual:.3e} vmax={st.v_max:.3f} "
              f"status={st.status} max_ff={st.max_ff_count} max_fs={st.max_fs_count} "
              f"counters={s.counters()}",
              file=sys.stderr, flush=True)
        N.check("wsph_step", rc)
    for k in range(args.warmup):
        N.check("wsph_step", lib.wsph_step(h, None))
    N.check("wsph_sync", lib.wsph_sync(h))
    c0 = s.counters()
    N.check("wsph_set_timing", lib.wsph_set_timing(h, 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        N.check("wsph_step", lib.wsph_step(h, None))
    N.check("wsph_sync", lib.wsph_sync(h))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    c1 = s.counters()
    tm = (C.c_double * 12)()
    N.check("wsph_last_timings", lib.wsph_last_timings(h, tm, 12))
    if c1["status"]:
        raise RuntimeError(f"device status {c1['status']} during the timed run")
    steps = c1["steps"] - c0["steps"]
    nonconv = c1["nonconverged"] - c0["nonconverged"]
    elapsed = t1 - t0
    h_rec = s.history(args.steps) if net is not None else None
    phases = phase_split(lib, h) if trace else None
    return dict(solver=s, n=s.n, elapsed=elapsed, steps=steps, tm=tm, phases=phases,
                history=h_rec,
                iters_per_step=(c1["iterations"] - c0["iterations"]) / max(1, steps),
                nonconverged=nonconv, ap_evals=c1["ap_evals"] - c0["ap_evals"],
                value=s.n * (steps - nonconv) / elapsed)


def bench_batch(args) -> None:
    """The batch workload: 256 independent dam-break scenes in one solver,
    launch shapes from the autotuner."""
    import torch
    from warpsph_amd import autotune, config, scene
    from warpsph_amd.batch import BatchIISPH
    start = {"zero": config.WARM_ZERO, "previous": config.WARM_PREVIOUS,
             "learned": config.WARM_LEARNED}[args.warm_start]
    net = None
    if start == config.WARM_LEARNED:
        from warpsph_amd.mlp import PressureNet
        net = PressureNet.load(args.net, max_batch=1024)
    cfgs = [c.replace(warm_start=start) for c in scene.batch_configs(config.preset(args.preset))]
    scs = [scene.build(c) for c in cfgs]

    def make(shape_cfg):
        return BatchIISPH([c.replace(**{k: getattr(shape_cfg, k) for k in autotune.SHAPE_KEYS})
                           for c in cfgs], scs, predictor=net)

    tuned = None
    shape_cfg = cfgs[0]
    if args.autotune != "off":
        tuned = autotune.tune(make, cfgs[0], key=f"{args.preset}|start{start}|n{sum(x.n for x in scs)}",
                              warmup=20,
                              force=args.autotune == "force",
                              log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
        shape_cfg = autotune.apply(cfgs[0], tuned["shape"])
    b = make(shape_cfg)
    lib, h = b._lib, b._h
    from warpsph_amd import _native as N
    for _ in range(args.warmup):
        N.check("wsph_step", lib.wsph_step(h, None))
    N.check("wsph_sync", lib.wsph_sync(h))
    c0 = b.counters()
    N.check("wsph_set_timing", lib.wsph_set_timing(h, 2))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        N.check("wsph_step", lib.wsph_step(h, None))
    N.check("wsph_sync", lib.wsph_sync(h))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    c1 = b.counters()
    tm = (C.c_double * 12)()
    N.check("wsph_last_timings", lib.wsph_last_timings(h, tm, 12))
    if c1["status"]:
        raise RuntimeError(f"device status {c1['status']} during the timed run")
    steps = c1["steps"] - c0["steps"]
    nonconv = c1["nonconverged"] - c0["nonconverged"]
    last = b.scene_stats()
    n = b.n
    ksteps = max(1, int(tm[11]))
    pair_ms = tm[10] / ksteps
    pair_bytes = n * BYTES_PER_AP_EVAL
    achieved = pair_bytes / (pair_ms * 1e-3) / 1e9 if tm[11] else 0.0
    line = {
        "metric": "fluid particle-steps/s, batch of independent scenes, rho-err<=0.1% per scene",
        "value": n * (steps - nonconv) / elapsed, "unit": "fluid particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.preset, "scenes": b.nscenes, "fluid_particles": n,
                   "boundary_particles": b.boundary_n, "spacing": cfgs[0].spacing,
                   "dt": cfgs[0].dt0, "warm_start": args.warm_start,
                   "shape": {k: getattr(shape_cfg, k) for k in autotune.SHAPE_KEYS}},
        "jacobi_iters_per_step": (c1["iterations"] - c0["iterations"]) / max(1, steps),
        "nonconverged_steps": nonconv,
        "last_step_scene_iterations": {"min": int(last["iterations"].min()),
                                       "median": float(np.median(last["iterations"])),
                                       "max": int(last["iterations"].max())},
        "phase_ms_per_step": phase_split(lib, h),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_j1_dev + k_j2_dev (one A p evaluation, first iteration)",
                     "algorithmic_bytes_per_launch": pair_bytes, "avg_launch_ms": pair_ms,
                     "launches_timed": int(tm[11])},
    }
    if tuned is not None:
        line["autotune"] = {"cached": tuned["cached"], "ms_per_step": tuned["ms_per_step"],
                            "candidates": len(tuned.get("table", []))}
    b.close()
    print(json.dumps(line), flush=True)


def bench_mesh(args) -> None:
    """Model sampling: lattice points inside a closed mesh per second."""
    import torch
    from warpsph_amd import _native as N
    from warpsph_amd import mesh as M
    lib = N.load()
    (v, t), axes = M.lattice_axes(MESH_SAMPLE, 1.0, MESH_SPACING)
    nx, ny, nz = (a.size for a in axes)
    npts, ncol, nt = nx * ny * nz, nx * ny, t.shape[0]
    dev = torch.device("cuda:0")
    dv, dtr = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (v, t))
    dx, dy, dz = (torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev) for a in axes)
    out = torch.zeros(npts, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def call():
        N.check("wsph_mesh_inside", lib.wsph_mesh_inside(
            dv.data_ptr(), v.shape[0], dtr.data_ptr(), nt, dx.data_ptr(), nx, dy.data_ptr(), ny,
            dz.data_ptr(), nz, out.data_ptr()))

    steps = max(1, min(args.steps, 50))
    for _ in range(args.warmup):
        call()
    ms = np.zeros(5)
    tm = (C.c_double * 7)()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
        N.check("wsph_mesh_timings", lib.wsph_mesh_timings(tm))
        ms += np.array(tm[:5])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms /= steps
    inside = int(out.sum().item())
    pairs, crossings = tm[5], tm[6]
    # algorithmic bytes per launch of each stage: every array it needs read
    # once, every output written once (MeshTri 128 B, index/count 4 B, height 8 B)
    stage_bytes = {
        "setup": nt * (12 + 3 * 24 + 128),
        "bin": nt * 128 + pairs * 4,
        "count_pass": pairs * (4 + 128) + ncol * 4,
        "fill_pass": pairs * (4 + 128) + ncol * 8 + crossings * 8,
        "classify": ncol * 8 + crossings * 8 * 2 + npts,
    }
    names = list(stage_bytes)
    dom = names[int(np.argmax(ms))]
    achieved = stage_bytes[dom] / (ms[names.index(dom)] * 1e-3) / 1e9
    line = {
        "metric": "lattice points classified inside a closed mesh per second (model sampling)",
        "value": npts * steps / elapsed, "unit": "points/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "mesh_sample", "mesh": "torus R 0.3 r 0.1, 192 x 96 quads",
                   "triangles": nt, "lattice": [nx, ny, nz], "spacing": MESH_SPACING,
                   "points_inside": inside},
        "kernel_ms": dict(zip(names, ms.tolist())),
        "triangle_tile_pairs": pairs, "column_crossings": crossings,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": {"setup": "k_mesh_setup", "bin": "k_mesh_bin",
                                "count_pass": "k_mesh_columns<false>",
                                "fill_pass": "k_mesh_columns<true>",
                                "classify": "k_mesh_sort + k_mesh_mark"}[dom],
                     "algorithmic_bytes_per_launch": stage_bytes[dom],
                     "avg_launch_ms": float(ms[names.index(dom)])},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        for threads, budget, key in ((1, args.cpu_seconds, "cpu_baseline"),
                                     (_nproc(), args.cpu_seconds / 2, "cpu_baseline_all_cores")):
            O.set_threads(threads)
            stride = 1
            while True:  # every stride-th column in x and z, all heights
                sub = [axes[0][::stride], axes[1], axes[2][::stride]]
                t1 = time.perf_counter()
                O.mesh_inside(v, t, *[a[:1] for a in sub[:2]] + [sub[2][:1]])
                per_point = time.perf_counter() - t1
                if per_point * sub[0].size * sub[1].size * sub[2].size / max(1, threads) <= budget:
                    break
                stride *= 2
            t1 = time.perf_counter()
            want = O.mesh_inside(v, t, *sub)
            el = time.perf_counter() - t1
            got = out.cpu().numpy().reshape(nx, ny, nz)[::stride, :, ::stride]
            line[key] = {"value": want.size / el, "unit": "points/s", "cores": threads,
                         "kind": "port", "host_cpu": _cpu_model(), "nproc": _nproc(),
                         "sample": f"every {stride}th lattice column in x and z "
                                   f"({want.size} points, {el:.1f} s), brute-force oracle",
                         "matches_gpu": bool(np.array_equal(got, want))}
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "warpsph_amd runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from warpsph_amd import config
    WARM.update(zero=config.WARM_ZERO, previous=config.WARM_PREVIOUS,
                learned=config.WARM_LEARNED)
    if args.preset == "mesh_sample":
        bench_mesh(args)
        return
    if config.preset(args.preset).scene == "batch":
        bench_batch(args)
        return

    base = config.preset(args.preset)
    base = base.replace(jacobi_mode={"persistent": config.JACOBI_PERSISTENT,
                                     "host": config.JACOBI_HOST,
                                     "device": config.JACOBI_DEVICE}[args.jacobi],
                        chunk_iters=args.chunk)
    if args.layout >= 0:
        base = base.replace(cell_layout=args.layout)
    if args.block:
        base = base.replace(block=args.block)
    if args.graph_iters >= 0:
        base = base.replace(graph_iters=args.graph_iters)
    if args.coop_per_cu:
        base = base.replace(coop_blocks_per_cu=args.coop_per_cu)
    net = None
    if args.warm_start == "learned":
        from warpsph_amd.mlp import PressureNet
        if not Path(args.net).exists():
            raise SystemExit(f"{args.net} is missing: train it with "
                             "`python -m warpsph_amd.train_predictor --out {args.net}`")
        net = PressureNet.load(args.net, max_batch=1024)
    tuned = None
    if args.autotune != "off":
        from warpsph_amd import autotune
        from warpsph_amd.solver import IISPH
        key = f"{args.preset}|start{WARM[args.warm_start]}|{args.jacobi}"
        tune_cfg = base.replace(warm_start=WARM[args.warm_start])
        cands = None
        if args.jacobi != "device":
            cands = [c for c in autotune.candidates() if c["jacobi_mode"] == tune_cfg.jacobi_mode]
        from warpsph_amd import scene as scn
        sc = scn.build(tune_cfg)
        tuned = autotune.tune(lambda c: IISPH(stage_path=None, config=c, scene=sc, predictor=net),
                              tune_cfg, warmup=20,
                              key=key, cands=cands, force=args.autotune == "force",
                              log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
        base = autotune.apply(base, tuned["shape"])
    cfg = base.replace(warm_start=WARM[args.warm_start])
    r = timed_run(cfg, net, args, trace=True)
    s, n, elapsed, steps, tm = r["solver"], r["n"], r["elapsed"], r["steps"], r["tm"]
    ap_evals = r["ap_evals"]
    ktm, kernel_source = tm, "the timed run (every 4th step)"
    if int(tm[11]) == 0:
        # graph replay or the persistent kernel: the first iteration's J1 and J2
        # run inside the captured graph or the one cooperative launch, where no
        # events are recorded; time the same kernels in a short eager run of
        # the same shape with the device-driven loop
        short = argparse.Namespace(**vars(args))
        short.steps, short.warmup = min(args.steps, 100), args.warmup
        eager = cfg.replace(graph_iters=0)
        if eager.jacobi_mode == config.JACOBI_PERSISTENT:
            eager = eager.replace(jacobi_mode=config.JACOBI_DEVICE, chunk_iters=1,
                                  coop_blocks_per_cu=0)
        q = timed_run(eager, net, short, trace=False)
        ktm = q["tm"]
        kernel_source = (f"an eager run of the same shape ({short.steps} steps)"
                         if eager.jacobi_mode == cfg.jacobi_mode else
                         f"an eager device-driven run of the same block shape ({short.steps} steps)")
        q["solver"].close()
    ksteps = max(1, int(ktm[11]))
    pair_ms = ktm[10] / ksteps
    pair_bytes = n * BYTES_PER_AP_EVAL
    achieved = pair_bytes / (pair_ms * 1e-3) / 1e9
    workload = args.preset
    line = {
        "metric": ("fluid particle-steps/s @1M, rho-err<=0.1%; Jacobi iters/step learned vs plain"
                   if args.preset == "dambreak_1m" else
                   f"fluid particle-steps/s, {args.preset}, rho-err<=0.1%"),
        "value": r["value"],
        "unit": "fluid particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "fluid_particles": n, "boundary_particles": s.boundary_n,
                   "spacing": cfg.spacing, "dt": cfg.dt0, "jacobi": args.jacobi,
                   "warm_start": args.warm_start, "eta": cfg.eta},
        "jacobi_iters_per_step": r["iters_per_step"],
        "nonconverged_steps": r["nonconverged"],
        "phase_ms_per_step": r["phases"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(workload)[0],
                     "traffic_source": traffic_from_profile(workload)[1],
                     "limiter": limiter_from_profile(workload),
                     "kernel": "k_j1_dev + k_j2_dev (one A p evaluation)",
                     "algorithmic_bytes_per_launch": pair_bytes,
                     "avg_launch_ms": pair_ms,
                     "events": "one before J1 and one after J2 of the first Jacobi iteration",
                     "launches_timed": int(ktm[11]), "timed_in": kernel_source},
        "step_roofline_frac": (n * (BYTES_PER_STEP_FIXED * steps + BYTES_PER_AP_EVAL * ap_evals)
                               / elapsed / 1e9) / HBM_PEAK_GBS,
    }
    line["config"]["shape"] = {k: getattr(cfg, k) for k in ("block", "jacobi_mode", "chunk_iters",
                                                            "coop_blocks_per_cu", "cell_layout",
                                                            "graph_iters")}
    if tuned is not None:
        line["autotune"] = {"cached": tuned["cached"], "ms_per_step": tuned["ms_per_step"],
                            "candidates": len(tuned.get("table", []))}
    if net is not None:
        h = r["history"]
        line["learned"] = {
            "steps_on_network": sum(1 for q in h if q["start_source"] == config.WARM_LEARNED),
            "start_err_learned_pa": sum(q["start_err_learned"] for q in h) / max(1, len(h)),
            "start_err_carry_pa": sum(q["start_err_carry"] for q in h) / max(1, len(h)),
            "params": str(Path(args.net).relative_to(ROOT)) if Path(args.net).is_relative_to(ROOT)
            else args.net}
    x_end = s.x.cpu().numpy()
    v_end = s.v.cpu().numpy()
    b_end = None
    if cfg.rigid_v is not None:
        b_end = (s.boundary_x.cpu().numpy(), s.boundary_v.cpu().numpy(), s.get_time()[0])
    s.close()
    plain = {}
    for name in [c for c in args.compare.split(",") if c and c != args.warm_start]:
        q = timed_run(base.replace(warm_start=WARM[name]), None, args, trace=False)
        plain[name] = {"value": q["value"], "ms_per_step": q["elapsed"] / args.steps * 1e3,
                       "jacobi_iters_per_step": q["iters_per_step"],
                       "nonconverged_steps": q["nonconverged"]}
        q["solver"].close()
        print(f"[bench] {name} start: {plain[name]}", file=sys.stderr, flush=True)
    if plain:
        line["plain_starts"] = plain
        if "previous" in plain:
            line["speedup_vs_previous_start"] = r["value"] / plain["previous"]["value"]
    if not args.no_cpu_baseline:
        ccfg = base.replace(warm_start=config.WARM_PREVIOUS)
        line["cpu_baseline"] = cpu_baseline(ccfg, x_end, v_end, args.cpu_seconds, boundary=b_end)
        line["cpu_baseline"]["host_cpu"] = _cpu_model()
        line["cpu_baseline"]["nproc"] = _nproc()
        # the same oracle on every core the process may use
        par = cpu_baseline(ccfg, x_end, v_end, args.cpu_seconds / 2, threads=_nproc(),
                           boundary=b_end)
        par["host_cpu"] = _cpu_model()
        par["nproc"] = _nproc()
        line["cpu_baseline_all_cores"] = par
        # the reference's own hash dims (IISPH.py:771-780): their aliased cells hold many
        # more candidates per query; one bounded sample on every core
        ref = cpu_baseline(ccfg.replace(grid="reference"), x_end, v_end, args.cpu_seconds / 2,
                           threads=_nproc(), boundary=b_end)
        ref["host_cpu"] = _cpu_model()
        ref["nproc"] = _nproc()
        line["cpu_reference_dims_all_cores"] = ref
    if not args.no_cpu_baseline and not args.no_cpu_faithful:
        # the reference's algorithm and hash dims on one thread, as Warp's CPU launches
        # run (one step, ~36 s at 1M particles)
        one = cpu_baseline(ccfg.replace(grid="reference"), x_end, v_end, 0.0, threads=1,
                           boundary=b_end)
        one["host_cpu"] = _cpu_model()
        one["nproc"] = _nproc()
        line["cpu_reference_faithful"] = one
    print(json.dumps(line), flush=True)


def _cpu_model():
    try:
        for l in open("/proc/cpuinfo"):
            if l.startswith("model name"):
                return l.split(":", 1)[1].strip() + f" x{os.cpu_count()}"
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
