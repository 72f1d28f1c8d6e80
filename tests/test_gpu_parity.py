"""Parity of the HIP path (libwsph.so through the C-ABI) with the CPU oracle.

Both sides run the same scene from the same float32 inputs. Because the GPU
kernels enumerate neighbours in the reference's order and keep its float32
operation order with no fused multiply-add, every field is compared for
exact equality (bit-exact), not within a tolerance.
"""
import sys

import numpy as np
import pytest

from helpers import assert_same, oracle_setup
from warpsph_amd import config, scene
from warpsph_amd import _native as N

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def gpu_solver(cfg, sc=None):
    from warpsph_amd.solver import IISPH
    return IISPH(stage_path=None, config=cfg, scene=sc, collect_stats=True)


def fields_of(s):
    return dict(rho=s.rho, v_adv=s.v_adv, rho_adv=s.rho_adv, d_ii=s.term_d_ii,
                a_ii=s.term_a_ii, p=s.p, Ap=s.term_Ap_i, s=s.sum_d_ij_p_j,
                ff_count=s.ff_neighbor_num, fs_count=s.fs_neighbor_num)


def compare_step(s, info, x_o, v_o, skip=()):
    f = {k: v.cpu().numpy() for k, v in fields_of(s).items()}
    for k in ("rho", "v_adv", "rho_adv", "d_ii", "a_ii", "p", "Ap", "s", "ff_count", "fs_count"):
        if k not in skip:
            assert_same(k, f[k], info[k])
    assert_same("x", s.x.cpu().numpy(), x_o)
    assert_same("v", s.v.cpu().numpy(), v_o)
    st = s.step_stats()
    assert st["iterations"] == info["iterations"]
    assert st["residual"] == info["residual"]


def test_boundary_psi_bitexact(orc):
    cfg = config.dambreak_8k()
    _announce("psi dambreak_8k")
    s = gpu_solver(cfg)
    sc, P, wt, gt = oracle_setup(cfg)
    assert_same("psi", s.boundary_phi.cpu().numpy(), orc.boundary_psi(P, wt, sc.xb))


@pytest.mark.parametrize("layout", [config.LAYOUT_MORTON, config.LAYOUT_LINEAR])
def test_neighbor_lists_bitexact(orc, layout):
    cfg = config.dambreak_8k().replace(cell_layout=layout, jitter=0.01)
    _announce(f"neighbour lists layout={layout}")
    s = gpu_solver(cfg)
    s.neighbor_search()
    sc, P, wt, gt = oracle_setup(cfg)
    for which, (xs, excl, dims) in enumerate([(sc.x, True, P.grid_dims),
                                              (sc.xb, False, P.bgrid_dims)]):
        cnt_o, idx_o, _ = orc.neighbors(sc.x, xs, cfg.h, tuple(dims), excl, cap=64)
        cnt, idx = s.neighbors(which, cap=64)
        cnt, idx = cnt.cpu().numpy(), idx.cpu().numpy()
        assert_same(f"count[{which}]", cnt, cnt_o)
        for i in range(sc.n):
            assert idx[i, :cnt[i]].tolist() == idx_o[i, :cnt_o[i]].tolist(), (which, i)


def test_cell_keys_bitexact(orc):
    cfg = config.dambreak_8k().replace(grid="reference", jitter=0.02)
    _announce("cell keys")
    s = gpu_solver(cfg)
    keys = s.field(N.F_CELL_KEY).cpu().numpy()
    sc = scene.build(cfg)
    assert_same("keys", keys, orc.hash_keys(sc.x, cfg.h, cfg.reference_grid_dims()))


@pytest.mark.parametrize("mode,chunk", [(config.JACOBI_PERSISTENT, 0), (config.JACOBI_HOST, 0),
                                        (config.JACOBI_DEVICE, 0), (config.JACOBI_DEVICE, 1),
                                        (config.JACOBI_DEVICE, 7)])
@pytest.mark.parametrize("force", [3, 0])
def test_one_step_bitexact(orc, mode, chunk, force):
    cfg = config.dambreak_8k().replace(jacobi_mode=mode, force_iters=force, chunk_iters=chunk)
    _announce(f"one step mode={mode} chunk={chunk} force={force}")
    s = gpu_solver(cfg)
    s.step()
    sc, P, wt, gt = oracle_setup(cfg)
    psi = orc.boundary_psi(P, wt, sc.xb)
    x_o, v_o, info = orc.step(P, wt, gt, sc.x, np.zeros_like(sc.x), sc.xb, sc.vb, psi)
    compare_step(s, info, x_o, v_o)


@pytest.mark.parametrize("preset,steps,mode", [("dambreak_8k", 25, config.JACOBI_DEVICE),
                                              ("dambreak_8k", 12, config.JACOBI_PERSISTENT)])
def test_trajectory_bitexact(orc, preset, steps, mode):
    cfg = config.preset(preset).replace(jacobi_mode=mode, warm_start=config.WARM_PREVIOUS)
    _announce(f"trajectory {preset} {steps} steps")
    s = gpu_solver(cfg)
    sc = scene.build(cfg)
    O = orc.OracleSolver(cfg, sc)
    for k in range(steps):
        s.step()
        info = O.step()
        assert s.step_stats()["iterations"] == info["iterations"], k
    assert_same("x", s.x.cpu().numpy(), O.x)
    assert_same("v", s.v.cpu().numpy(), O.v)
    assert_same("p", s.p.cpu().numpy(), O.p)


def test_warm_start_previous_bitexact(orc):
    cfg = config.dambreak_8k().replace(warm_start=config.WARM_PREVIOUS)
    _announce("warm start previous")
    s = gpu_solver(cfg)
    sc = scene.build(cfg)
    O = orc.OracleSolver(cfg, sc)
    for k in range(4):
        s.step()
        info = O.step()
        assert s.step_stats()["iterations"] == info["iterations"], k
        assert s.step_stats()["ap_evals"] == info["ap_evals"], k
    compare_step(s, info, O.x, O.v)


def test_external_start_bitexact(orc):
    cfg = config.dambreak_8k().replace(warm_start=config.WARM_EXTERNAL)
    _announce("external start")
    s = gpu_solver(cfg)
    sc, P, wt, gt = oracle_setup(cfg)
    psi = orc.boundary_psi(P, wt, sc.xb)
    p0 = np.random.default_rng(5).uniform(0, 2000, sc.n).astype(np.float32)
    s.set_pressure(p0)
    s.step()
    x_o, v_o, info = orc.step(P, wt, gt, sc.x, np.zeros_like(sc.x), sc.xb, sc.vb, psi, p0=p0)
    compare_step(s, info, x_o, v_o)


def test_reference_scene_one_step(orc):
    """The reference's default scene: 102,641 fluid and 124,056 wall
    particles, CFL dt, 8192/65536-sample tables (global-memory table path)."""
    cfg = config.reference().replace(force_iters=2)
    _announce("reference scene")
    s = gpu_solver(cfg)
    assert (s.n, s.boundary_n) == (102641, 124056)
    s.step()
    sc, P, wt, gt = oracle_setup(cfg)
    psi = orc.boundary_psi(P, wt, sc.xb)
    x_o, v_o, info = orc.step(P, wt, gt, sc.x, np.zeros_like(sc.x), sc.xb, sc.vb, psi)
    compare_step(s, info, x_o, v_o)
    t, dt = s.get_time()
    assert t == cfg.dt_max
    assert dt == min(cfg.cfl * cfg.spacing / max(float(info["v_max"]), cfg.inv_small), cfg.dt_max)


def test_reference_scene_trajectory(orc):
    """The reference's own configuration over several steps: its default
    scene, p0 = 0, the full Jacobi loop and the CFL clock (IISPH.py:1367-1371;
    dt stays at TIME_STEP_MAX while v_max < 8 m/s). The device's clock, the
    particle state and every step's iteration count follow the oracle's bit
    for bit."""
    cfg = config.reference()
    _announce("reference scene, 6 steps")
    s = gpu_solver(cfg)
    sc = scene.build(cfg)
    O = orc.OracleSolver(cfg, sc)
    for k in range(6):
        s.step()
        info = O.step()
        st = s.step_stats()
        assert st["iterations"] == info["iterations"], k
        assert st["residual"] == info["residual"], k
        assert s.get_time() == (O.sim_time, O.dt), k
    assert_same("x", s.x.cpu().numpy(), O.x)
    assert_same("v", s.v.cpu().numpy(), O.v)
    assert_same("p", s.p.cpu().numpy(), O.p)


def test_cfl_dt_trajectory(orc):
    """dt = min(cfl * d / max(v_max, INV_SMALL), TIME_STEP_MAX) each step
    (IISPH.py:1368-1370) on the reference's scene at d = 0.1 with a CFL
    number small enough that the limit binds from the first step and dt
    changes every step: clock, state and iteration counts bit-identical to
    the oracle's."""
    cfg = config.reference().replace(spacing=0.1, cfl=0.002, warm_start=config.WARM_PREVIOUS)
    _announce("CFL dt, 15 steps")
    s = gpu_solver(cfg)
    sc = scene.build(cfg)
    O = orc.OracleSolver(cfg, sc)
    dts = []
    for k in range(15):
        s.step()
        info = O.step()
        assert s.step_stats()["iterations"] == info["iterations"], k
        assert s.get_time() == (O.sim_time, O.dt), k
        dts.append(O.dt)
    assert len(set(dts)) > 10 and min(dts) < cfg.dt_max
    assert_same("x", s.x.cpu().numpy(), O.x)
    assert_same("p", s.p.cpu().numpy(), O.p)


def test_viscosity_term_bitexact(orc):
    """The Laplacian viscosity term as written (IISPH.py:249-278) with VIS_MU != 0."""
    cfg = config.dambreak_8k().replace(vis_mu=0.5, xsph=0.0, force_iters=2)
    _announce("viscosity")
    sc = scene.build(cfg)
    v0 = np.random.default_rng(9).normal(0, 0.1, sc.x.shape).astype(np.float32)
    s = gpu_solver(cfg, sc)
    import ctypes
    import torch
    vt = torch.from_numpy(v0).cuda()
    xt = torch.from_numpy(sc.x).cuda()
    torch.cuda.synchronize()
    N.check("wsph_set_fluid", s._lib.wsph_set_fluid(s._h, xt.data_ptr(), vt.data_ptr(), sc.n))
    s.step()
    _, P, wt, gt = oracle_setup(cfg, sc)
    psi = orc.boundary_psi(P, wt, sc.xb)
    x_o, v_o, info = orc.step(P, wt, gt, sc.x, v0, sc.xb, sc.vb, psi)
    compare_step(s, info, x_o, v_o)


def test_drivers_agree_at_full_size():
    """At the 1M-particle size the three Jacobi drivers (host test per
    iteration, persistent kernel, device-side stop test) give identical
    states: the arithmetic is the same and the loop stops at the same
    iteration."""
    cfg = config.dambreak_1m().replace(warm_start=config.WARM_PREVIOUS)
    out = {}
    for mode in (config.JACOBI_DEVICE, config.JACOBI_PERSISTENT, config.JACOBI_HOST):
        _announce(f"1M drivers mode={mode}")
        s = gpu_solver(cfg.replace(jacobi_mode=mode))
        its = []
        for _ in range(3):
            s.step()
            its.append(s.step_stats()["iterations"])
        out[mode] = (its, s.x.cpu().numpy(), s.v.cpu().numpy(), s.p.cpu().numpy())
        s.close()
    ref = out[config.JACOBI_DEVICE]
    for mode, got in out.items():
        assert got[0] == ref[0], (mode, got[0], ref[0])
        for a, b in zip(got[1:], ref[1:]):
            assert_same(f"mode {mode}", a, b)


SHAPES = [dict(block=b, cell_layout=l, chunk_iters=ch, graph_iters=g)
          for b, l, ch, g in ((128, config.LAYOUT_MORTON, 1, 0), (256, config.LAYOUT_LINEAR, 2, 0),
                              (512, config.LAYOUT_MORTON, 1, 0), (512, config.LAYOUT_LINEAR, 4, 0),
                              (512, config.LAYOUT_MORTON, 2, 2), (1024, config.LAYOUT_MORTON, 1, 0))]


@pytest.mark.parametrize("warm", [config.WARM_PREVIOUS, config.WARM_LEARNED])
def test_launch_shapes_agree(orc, warm):
    """The launch shapes the autotuner chooses from (block 128 to 1024, Morton
    or linear cells, device chunks of 1, 2 or 4 iterations, a captured graph)
    give the same states bit for bit: the J2 LDS reservation, the list ring
    per wave and the XCD tile mapping depend on the block size, the results
    may not. With the previous-step start every shape also equals the oracle;
    with the learned start (whose network runs only on the device) they equal
    each other, start sources and iteration counts included."""
    from warpsph_amd.mlp import PressureNet
    from warpsph_amd.solver import IISPH
    cfg = config.dambreak_8k().replace(jitter=0.01, warm_start=warm)
    sc = scene.build(cfg)
    net = None
    if warm == config.WARM_LEARNED:
        net = PressureNet.load(str(__import__("pathlib").Path(__file__).resolve().parents[1]
                                   / "warpsph_amd" / "pressure_net.npy"), max_batch=1024)
    o = orc.OracleSolver(cfg, sc) if warm == config.WARM_PREVIOUS else None
    ref = None
    for shape in SHAPES:
        _announce(f"launch shape {shape}, start {warm}")
        s = IISPH(stage_path=None, config=cfg.replace(**shape), scene=sc, collect_stats=True,
                  predictor=net)
        its = []
        for _ in range(12):
            s.step()
            its.append(s.step_stats()["iterations"])
        got = (its, [r["start_source"] for r in s.history()], s.x.cpu().numpy(),
               s.v.cpu().numpy(), s.p.cpu().numpy())
        s.close()
        if ref is None:
            ref = got
            if o is not None:
                infos = [o.step() for _ in range(12)]
                assert its == [q["iterations"] for q in infos]
                assert_same("x vs oracle", got[2], o.x)
                assert_same("v vs oracle", got[3], o.v)
                assert_same("p vs oracle", got[4], infos[-1]["p"])
        assert got[0] == ref[0] and got[1] == ref[1], (shape, got[0], ref[0])
        for name, a_, b_ in zip(("x", "v", "p"), got[2:], ref[2:]):
            assert_same(f"{name} {shape}", a_, b_)


@pytest.mark.parametrize("warm", [config.WARM_ZERO, config.WARM_PREVIOUS])
def test_dambreak_200_step_statistics(orc, warm):
    """200 steps of the 8,000-particle dam break: per-step Jacobi iteration
    counts and residuals equal the oracle's, and the statistics (mean density
    error, kinetic energy, free-surface and mean height) agree to well under
    the 1% bar (the trajectories are bit-identical; only the float64
    summation order of the statistics differs)."""
    cfg = config.dambreak_8k().replace(warm_start=warm)
    _announce(f"200-step statistics warm={warm}")
    s = gpu_solver(cfg)
    sc = scene.build(cfg)
    O = orc.OracleSolver(cfg, sc)
    ref = []
    for k in range(200):
        s.step()
        info = O.step()
        v = O.v.astype(np.float64)
        ref.append(dict(iterations=info["iterations"], residual=info["residual"],
                        kinetic_energy=0.5 * float(np.float32(cfg.mass)) * float((v * v).sum()),
                        surface_height=float(O.x[:, 1].max()),
                        mean_height=float(O.x[:, 1].astype(np.float64).mean())))
    hist = s.history()[-200:]
    assert len(hist) == 200
    for k, (g, r) in enumerate(zip(hist, ref)):
        assert g["iterations"] == r["iterations"], k
        assert g["residual"] == r["residual"], k
        assert g["surface_height"] == r["surface_height"], k
        for key in ("kinetic_energy", "mean_height"):
            assert abs(g[key] - r[key]) <= 1e-9 * abs(r[key]), (k, key, g[key], r[key])
    assert_same("x", s.x.cpu().numpy(), O.x)
