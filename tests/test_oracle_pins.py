"""Pins of the oracle's predict_v_adv (XSPH and the Laplacian viscosity term),
predict_rho_adv, update_p and the learned start's features and arbiter:
values the mathematics fixes, computed here in float64 from the closed-form
cubic spline (not from the tables the oracle interpolates), on hand-built
scenes of one to three particles, interior lattices and a converged Jacobi
solve.

Each ``pin_*`` function raises AssertionError when the oracle disagrees;
tests/test_oracle_mutants.py builds copies of the oracle with plausible
mistakes planted (a sign flip on v_b, v in place of v_adv, a dropped psi_b or
omega, sign(0) = -1, no clamp, ...) and checks that the pins catch each one.

Reference: wp_sph/solver/IISPH.py:217-278 (predict_v_adv), :281-329
(predict_rho_adv), :521-539 (update_p); the learned start has no reference
counterpart (oracle/learn_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import kernel_dw, kernel_grad, kernel_sigma, kernel_w
from warpsph_amd import config, scene, tables
from warpsph_amd._native import make_params

DIMS = (8, 8, 8)  # hash dims >= 3 per axis: any such dims give the same lists


def base_cfg(**kw):
    """C0's physics (d = 0.02, h = 0.04, m = 8e-3, dt = 1e-3 fixed)."""
    return config.dambreak_8k().replace(**kw)


def tiny_step(O, cfg, x, v, xb=None, vb=None):
    """One step of a hand-built scene through ``O``: the oracle module, or any
    object with its ``boundary_psi`` and ``step`` signatures (the GPU tests
    pass one that runs libwsph.so). psi comes from ``O.boundary_psi`` and is
    returned in the step's fields; a lone boundary particle has
    psi = rho0 / W(0)."""
    x = np.asarray(x, np.float32).reshape(-1, 3)
    v = np.asarray(v, np.float32).reshape(-1, 3)
    xb = np.zeros((0, 3), np.float32) if xb is None else np.asarray(xb, np.float32).reshape(-1, 3)
    vb = np.zeros_like(xb) if vb is None else np.asarray(vb, np.float32).reshape(-1, 3)
    P = make_params(cfg, DIMS, DIMS)
    wt, gt = tables.make_tables(cfg.h, cfg.w_size, cfg.grad_size)
    psi = O.boundary_psi(P, wt, xb) if xb.shape[0] else np.zeros(0, np.float32)
    _, _, info = O.step(P, wt, gt, x, v, xb, vb, psi)
    info["psi"] = psi
    return info


def f32(a):
    return np.asarray(a, np.float32)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def close(got, want, rtol, what):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    err = np.max(np.abs(got - want))
    scale = max(np.max(np.abs(want)), 1e-300)
    assert err <= rtol * scale, f"{what}: got {got}, closed form {want} (rel err {err / scale:.2e})"


def lattice(cells, d, origin=(0.1, 0.1, 0.1)):
    return scene.lattice_block(cells, d, origin).astype(np.float32)


def center_index(cells):
    return ((cells[0] // 2) * cells[1] + cells[1] // 2) * cells[2] + cells[2] // 2


# ---- predict_rho_adv (IISPH.py:281-329) --------------------------------------
def pin_rho_adv_uniform(O):
    """A uniform velocity field without boundary: v_adv = v + g dt exactly
    (the XSPH sum of equal velocities is zero) and rho_adv = rho exactly
    (every v_adv,i - v_adv,j is zero)."""
    cfg = base_cfg()
    x = lattice((6, 6, 6), cfg.spacing)
    v = np.tile(f32([0.3, -0.2, 0.1]), (x.shape[0], 1))
    info = tiny_step(O, cfg, x, v)
    want_va = v + f32([0.0, -cfg.gravity, 0.0]) * np.float32(cfg.dt)
    assert np.array_equal(info["v_adv"], want_va), "v_adv of a uniform field is v + g dt"
    assert np.array_equal(info["rho_adv"], info["rho"]), "rho_adv of a uniform field is rho"


def lattice_div_sum(d, h):
    """sum_j |W'(r_j)| r_jx^2 / r_j over the neighbours of an interior point
    of the cubic lattice with h = 2d: 2 axis neighbours at d, 8 face
    diagonals at sqrt2 d and 8 body diagonals at sqrt3 d have r_x^2 = d^2."""
    return (2 * kernel_dw(d, h) * d + 8 * kernel_dw(math.sqrt(2) * d, h) * d / math.sqrt(2)
            + 8 * kernel_dw(math.sqrt(3) * d, h) * d / math.sqrt(3))


def pin_rho_adv_linear_field(O):
    """v_x = alpha (x - x_c) on an interior lattice, g = 0, no XSPH:
    rho_adv - rho = dt m sum_j (v_i - v_j) . grad W_ij
                  = -alpha dt m sum_j |W'(r_j)| r_jx^2 / r_j,
    the lattice form of -rho dt div v (m times the lattice sum is 1.02 rho0
    at h = 2d)."""
    cfg = base_cfg(gravity=0.0, xsph=0.0)
    cells = (9, 9, 9)
    x = lattice(cells, cfg.spacing)
    alpha = 20.0
    c = center_index(cells)
    v = np.zeros_like(x)
    v[:, 0] = alpha * (x[:, 0].astype(np.float64) - float(x[c, 0]))
    info = tiny_step(O, cfg, x, v)
    k = cfg.mass * lattice_div_sum(cfg.spacing, cfg.h) / cfg.rho0
    assert abs(k - 1.0201) < 1e-3, k
    want = -alpha * cfg.dt * cfg.mass * lattice_div_sum(cfg.spacing, cfg.h)
    close(float(info["rho_adv"][c]) - float(info["rho"][c]), want, 2e-4,
          "rho_adv - rho of a linear field")


def pin_rho_adv_boundary_pair(O):
    """One fluid particle and one boundary particle at r = 0.6 h along e:
    rho = m W(0) + psi W(r) and
    rho_adv - rho = dt psi (v_adv,i - v_b) . grad W_ib,
    with v_adv,i = v_i + g dt and grad W_ib = |W'(r)| e."""
    cfg = base_cfg()
    e = unit([0.48, -0.6, 0.64])
    xi = f32([0.5, 0.5, 0.5])
    xb = f32(xi + 0.6 * cfg.h * e)
    vi = f32([0.2, -0.3, 0.1])
    vb = f32([-0.1, 0.05, 0.2])
    info = tiny_step(O, cfg, [xi], [vi], [xb], [vb])
    r = float(np.linalg.norm(xb.astype(np.float64) - xi))
    psi32 = float(info["psi"][0])
    close(psi32, cfg.rho0 / kernel_w(0.0, cfg.h), 1e-6, "psi of a lone boundary particle")
    rho = cfg.mass * kernel_w(0.0, cfg.h) + psi32 * kernel_w(r, cfg.h)
    close(info["rho"][0], rho, 1e-5, "rho with one boundary neighbour")
    va = vi.astype(np.float64) + np.array([0.0, -cfg.gravity * cfg.dt, 0.0])
    close(info["v_adv"][0], va, 1e-6, "v_adv without fluid neighbours")
    gw = kernel_grad(xi, xb, cfg.h)
    want = cfg.dt * psi32 * np.dot(va - vb, gw)
    close(float(info["rho_adv"][0]) - float(info["rho"][0]), want, 1e-4,
          "rho_adv - rho with one boundary neighbour")


# ---- predict_v_adv: XSPH smoothing and the Laplacian term (IISPH.py:217-278) --
def line_of_three(cfg, r_over_h=0.6):
    e = unit([0.6, 0.0, 0.8])
    x0 = np.array([0.3, 0.4, 0.5])
    return f32([x0, x0 + r_over_h * cfg.h * e, x0 + 2 * r_over_h * cfg.h * e])


def pin_xsph_three_particles(O):
    """Three fluid particles on a line, neighbours 0-1 and 1-2 only:
    v_adv,i = v_i + g dt + c sum_j (v_j - v_i) (m / rho_j) W_ij with
    rho_0 = rho_2 = m (W(0) + W(r)) and rho_1 = m (W(0) + 2 W(r)); the
    coefficient uses the neighbour's density."""
    c = 0.2
    cfg = base_cfg(xsph=c)
    x = line_of_three(cfg)
    v = f32([[0.5, -0.2, 0.1], [-0.3, 0.4, 0.2], [0.1, 0.1, -0.6]])
    info = tiny_step(O, cfg, x, v)
    X = x.astype(np.float64)
    nb = {0: [1], 1: [0, 2], 2: [1]}
    w = {(i, j): kernel_w(float(np.linalg.norm(X[j] - X[i])), cfg.h) for i in nb for j in nb[i]}
    rho = np.array([cfg.mass * (kernel_w(0.0, cfg.h) + sum(w[(i, j)] for j in nb[i]))
                    for i in range(3)])
    close(info["rho"], rho, 1e-5, "rho of three particles in a line")
    g = np.array([0.0, -cfg.gravity * cfg.dt, 0.0])
    for i in range(3):
        want = c * sum((v[j].astype(np.float64) - v[i]) * cfg.mass / rho[j] * w[(i, j)]
                       for j in nb[i])
        got = info["v_adv"][i].astype(np.float64) - v[i] - g
        close(got, want, 2e-4, f"XSPH term of particle {i}")


def lattice_xsph_sum(d, h):
    """sum_j dx_j^2 W(r_j) / sum_j W(r_j) (self included in the
    denominator) over an interior point of the cubic lattice with h = 2d."""
    num = d * d * (2 * kernel_w(d, h) + 8 * kernel_w(math.sqrt(2) * d, h)
                   + 8 * kernel_w(math.sqrt(3) * d, h))
    den = (kernel_w(0.0, h) + 6 * kernel_w(d, h) + 12 * kernel_w(math.sqrt(2) * d, h)
           + 8 * kernel_w(math.sqrt(3) * d, h))
    return num / den


def pin_xsph_lattice_fields(O):
    """Interior lattice, g = 0: for a linear field the XSPH sum cancels by
    symmetry (v_adv = v to rounding); for v_x = alpha x^2 it is the uniform
    c alpha sum_j dx_j^2 (m / rho_j) W_ij = c alpha d^2 (2 W(d) + 8 W(sqrt2 d)
    + 8 W(sqrt3 d)) / sum W, since m / rho_j = 1 / sum W in the interior."""
    c = 0.2
    cfg = base_cfg(xsph=c, gravity=0.0)
    cells = (9, 9, 9)
    x = lattice(cells, cfg.spacing)
    k = center_index(cells)
    X = x.astype(np.float64)
    alpha = 100.0
    v = np.zeros_like(x)
    v[:, 0] = 0.5 * (X[:, 0] - X[k, 0])
    info = tiny_step(O, cfg, x, v)
    assert abs(float(info["v_adv"][k, 0]) - float(v[k, 0])) < 1e-7, "linear field: XSPH cancels"
    v[:, 0] = alpha * (X[:, 0] - X[k, 0]) ** 2
    info = tiny_step(O, cfg, x, v)
    want = c * alpha * lattice_xsph_sum(cfg.spacing, cfg.h)
    close(float(info["v_adv"][k, 0]) - float(v[k, 0]), want, 1e-3, "XSPH of a quadratic field")


def pin_laplacian_uniform(O):
    """VIS_MU != 0, uniform v, no boundary: every (v_i - v_j) . x_ij is zero,
    so v_adv = v + g dt exactly."""
    cfg = base_cfg(vis_mu=0.5, xsph=0.0)
    x = lattice((6, 6, 6), cfg.spacing)
    v = np.tile(f32([0.3, -0.2, 0.1]), (x.shape[0], 1))
    info = tiny_step(O, cfg, x, v)
    want = v + f32([0.0, -cfg.gravity, 0.0]) * np.float32(cfg.dt)
    assert np.array_equal(info["v_adv"], want), "Laplacian term of a uniform field is zero"


def pin_laplacian_three_particles(O):
    """Three fluid particles on a line, neighbours 0-1 and 1-2 only
    (IISPH.py:249-260, 273-278):
    L_i = 10 m sum_j -grad W_ij ((v_i - v_j) . x_ij) / rho_j / r^2
        = -10 m sum_j |W'(r)| ((v_i - v_j) . e_ij) e_ij / (rho_j r),
    v_adv,i = v_i + (VIS_MU L_i / rho_i + g) dt, with rho_0 = rho_2 =
    m (W(0) + W(r)) and rho_1 = m (W(0) + 2 W(r)), so the neighbour's density
    and the particle's own one are told apart."""
    mu = 0.5
    cfg = base_cfg(vis_mu=mu, xsph=0.0)
    x = line_of_three(cfg, 0.55)
    v = f32([[0.4, -0.1, 0.3], [-0.2, 0.35, -0.15], [0.1, 0.25, 0.5]])
    info = tiny_step(O, cfg, x, v)
    X = x.astype(np.float64)
    nb = {0: [1], 1: [0, 2], 2: [1]}
    rho = np.array([cfg.mass * (kernel_w(0.0, cfg.h) + sum(
        kernel_w(float(np.linalg.norm(X[j] - X[i])), cfg.h) for j in nb[i])) for i in range(3)])
    g = np.array([0.0, -cfg.gravity * cfg.dt, 0.0])
    for i in range(3):
        lap = np.zeros(3)
        for j in nb[i]:
            r = float(np.linalg.norm(X[j] - X[i]))
            ee = (X[j] - X[i]) / r
            lap -= 10 * cfg.mass * kernel_dw(r, cfg.h) * np.dot(v[i].astype(np.float64) - v[j],
                                                                 ee) * ee / (rho[j] * r)
        want = mu * lap / rho[i] * cfg.dt
        close(info["v_adv"][i].astype(np.float64) - v[i] - g, want, 2e-4,
              f"Laplacian term of particle {i}")


def pin_laplacian_boundary_pair(O):
    """One fluid and one boundary particle at r = 0.55 h along e
    (IISPH.py:262-278): L_i = 10 / rho0 sum_b -grad W_ib psi_b (v_i . x_ib) / r^2
    = -10 psi |W'(r)| (v_i . e) e / (rho0 r); the boundary velocity does not
    enter."""
    mu = 0.5
    cfg = base_cfg(vis_mu=mu, xsph=0.0)
    e = unit([-0.64, 0.0, 0.768])
    xi = f32([0.4, 0.3, 0.6])
    xb = f32(xi + 0.55 * cfg.h * e)
    vi = f32([0.3, -0.25, 0.45])
    info = tiny_step(O, cfg, [xi], [vi], [xb], [[0.7, 0.0, -0.2]])
    psi32 = float(info["psi"][0])
    r = float(np.linalg.norm(xb.astype(np.float64) - xi))
    ee = (xb.astype(np.float64) - xi) / r
    rho = cfg.mass * kernel_w(0.0, cfg.h) + psi32 * kernel_w(r, cfg.h)
    lap = -10 * psi32 * kernel_dw(r, cfg.h) * np.dot(vi.astype(np.float64), ee) * ee \
        / (cfg.rho0 * r)
    want = mu * lap / rho * cfg.dt
    g = np.array([0.0, -cfg.gravity * cfg.dt, 0.0])
    close(info["v_adv"][0].astype(np.float64) - vi - g, want, 2e-4,
          "Laplacian term with one boundary neighbour")


# ---- update_p (IISPH.py:521-539) inside the Jacobi loop ------------------------
def pin_update_p_isolated(O):
    """An isolated fluid particle: a_ii = 0, so the guard uses sign(0) * 1e-6
    = +1e-6 (Warp's sign(0) is +1); A p = 0 and rho_adv = rho = m W(0) < rho0,
    so the residual set is empty (e = 0) and the loop stops after its
    minimum of 2 iterations: p = 2 omega (rho0 - m W(0)) / 1e-6."""
    cfg = base_cfg()
    info = tiny_step(O, cfg, [[0.5, 0.5, 0.5]], [[0.0, 0.0, 0.0]])
    assert info["a_ii"][0] == 0.0
    assert info["iterations"] == 2
    want = 2 * cfg.omega * (cfg.rho0 - cfg.mass * kernel_w(0.0, cfg.h)) / cfg.inv_small
    close(info["p"][0], want, 1e-5, "p of an isolated particle after two updates")


def pin_update_p_guard_and_clamp(O):
    """Two fluid particles at 0.99 h: a_ii = -2 m^2 |W'(r)|^2 dt^2 / rho^2 lies
    inside the guard band (-1e-6, 0), so the divisor is -1e-6; with
    rho_adv < rho0 the update omega (rho0 - rho_adv - A p) / (-1e-6) is
    negative and the clamp leaves p = 0."""
    cfg = base_cfg(gravity=0.0, xsph=0.0)
    e = unit([1.0, 0.2, -0.3])
    xi = f32([0.5, 0.5, 0.5])
    x = [xi, f32(xi + 0.99 * cfg.h * e)]
    info = tiny_step(O, cfg, x, np.zeros((2, 3), np.float32))
    r = float(np.linalg.norm(np.float64(x[1]) - x[0]))
    rho = cfg.mass * (kernel_w(0.0, cfg.h) + kernel_w(r, cfg.h))
    a = -2 * cfg.mass ** 2 * kernel_dw(r, cfg.h) ** 2 * cfg.dt ** 2 / rho ** 2
    close(info["a_ii"], [a, a], 1e-3, "a_ii of a pair")
    assert -cfg.inv_small < a < 0.0
    assert np.all(info["rho_adv"] < cfg.rho0)
    assert np.all(info["p"] == 0.0), f"clamped pressure, got {info['p']}"


def pin_jacobi_fixed_point(O):
    """Relaxed Jacobi (eta = 0) on a small dam break converges to the
    conditions of the clamped problem: rho_adv + A p = rho0 where p > 0, and
    rho_adv + A p <= rho0 where p = 0 (a particle the update would push below
    zero); the largest violation shrinks geometrically with the iterations."""
    cfg = base_cfg(fluid_cells=(8, 8, 8))
    sc = scene.build(cfg)
    gd, bd = scene.grid_dims_for(cfg, sc)
    P = make_params(cfg, gd, bd)
    wt, gt = tables.make_tables(cfg.h, cfg.w_size, cfg.grad_size)
    psi = O.boundary_psi(P, wt, sc.xb)

    def solve(eta, force):
        P.eta, P.force_iters = eta, force
        _, _, info = O.step(P, wt, gt, sc.x, np.zeros_like(sc.x), sc.xb, sc.vb, psi)
        return info

    def violation(info):
        rs = (info["rho_adv"].astype(np.float64) + info["Ap"]) / cfg.rho0 - 1.0
        pos = info["p"] > 0
        strong = np.abs(info["a_ii"]) >= cfg.inv_small
        return max(np.max(np.abs(rs[pos])), np.max(rs[~pos & strong], initial=0.0)), pos.mean()

    full = solve(0.0, 0)
    assert full["iterations"] == 101
    v101, frac = violation(full)
    v2, _ = violation(solve(0.0, 2))
    v30, _ = violation(solve(0.0, 30))
    assert frac > 0.3, frac  # the clamp leaves about half the block compressed
    # geometric convergence: about 0.087, 0.0075 and 0.00034 after 2, 30 and 101
    assert v101 < 1e-3, v101
    assert v101 < 0.1 * v30 < 0.01 * v2, (v2, v30, v101)


# ---- learned start: features f4-f7 and the arbiter (oracle/learn_oracle.py) ---
def pin_learned_lattice_constant(L):
    """a_ref's lattice sum of |grad W|^2 at h = 2d equals
    (sigma / h)^2 4 pi^2 0.42483995, the constant of a_ii's closed form."""
    for d in (0.02, 0.01):
        h = 2 * d
        got = L.lattice_grad_sq_sum(h, d)
        want = (kernel_sigma(h) / h) ** 2 * 4 * math.pi ** 2 * 0.42483995
        close(got, want, 1e-7, "lattice sum of |grad W|^2")


def pin_learned_means(L, cfg=None):
    """Columns 4-7 are sum_j W_ij f_j / sum_j W_ij over the particle and its
    fluid neighbours: a constant column gives the constant; for two
    particles at r, (W(0) f_i + W(r) f_j) / (W(0) + W(r))."""
    cfg = cfg or base_cfg()
    wt, _ = tables.make_tables(cfg.h, cfg.w_size, cfg.grad_size)
    dt = cfg.dt
    x = lattice((4, 4, 4), cfg.spacing)
    n = x.shape[0]
    nbr = np.full((n, 64), -1, np.int32)
    cnt = np.zeros(n, np.int32)
    for i in range(n):
        j = [k for k in range(n) if k != i and np.linalg.norm(x[k] - x[i]) < cfg.h]
        cnt[i] = len(j)
        nbr[i, :len(j)] = j
    a_ref = L.a_ref(cfg, dt)
    f = L.features(cfg, dt, x, np.full(n, 1003.0), np.full(n, 0.7 * a_ref), np.full(n, 24),
                   np.full(n, 5000.0), nbr, wt)
    close(f[:, 4:], f[:, :4], 1e-12, "means of constant features")
    close(f[0, :4], [0.3, 0.7, 0.75, 5000.0 / L.p_ref(cfg)], 1e-12, "self features")
    e = unit([0.2, 0.9, -0.3])
    x2 = f32([[0.5, 0.5, 0.5], 0.5 + 0.45 * cfg.h * e])
    nbr2 = np.array([[1], [0]], np.int32)
    f2 = L.features(cfg, dt, x2, np.array([1010.0, 995.0]), np.array([a_ref, 0.25 * a_ref]),
                    np.array([1, 1]), np.array([100.0, 900.0]), nbr2, wt)
    r = float(np.linalg.norm(x2[1].astype(np.float64) - x2[0]))
    w0, w1 = kernel_w(0.0, cfg.h), kernel_w(r, cfg.h)
    for i, j in ((0, 1), (1, 0)):
        close(f2[i, 4:], (w0 * f2[i, :4] + w1 * f2[j, :4]) / (w0 + w1), 1e-6,
              f"neighbour means of particle {i}")


def pin_learned_arbiter(L):
    """Two particles with a_ii = 2 a_ref and 0.5 a_ref (weights 1 and 0.5)
    and start errors 3 and -4 Pa: sqrt(((1*3)^2 + (0.5*4)^2) / 2)."""
    cfg = base_cfg()
    a = L.a_ref(cfg, cfg.dt)
    got = L.arbiter_error([103.0, 46.0], [100.0, 50.0], [2 * a, 0.5 * a], cfg, cfg.dt)
    close(got, math.sqrt(6.5), 1e-12, "arbiter error")


ORACLE_PINS = [pin_rho_adv_uniform, pin_rho_adv_linear_field, pin_rho_adv_boundary_pair,
               pin_xsph_three_particles, pin_xsph_lattice_fields, pin_laplacian_uniform,
               pin_laplacian_three_particles, pin_laplacian_boundary_pair, pin_update_p_isolated,
               pin_update_p_guard_and_clamp, pin_jacobi_fixed_point]
LEARN_PINS = [pin_learned_lattice_constant, pin_learned_means, pin_learned_arbiter]


@pytest.mark.parametrize("pin", ORACLE_PINS, ids=lambda f: f.__name__)
def test_oracle_pin(orc, pin):
    pin(orc)


@pytest.mark.parametrize("pin", LEARN_PINS, ids=lambda f: f.__name__)
def test_learn_oracle_pin(pin):
    import learn_oracle as L
    pin(L)
