"""Pins the CPU restatement of the pressure predictor (oracle/mlp_oracle.py):
its gradient against central finite differences and against torch autograd,
its bfloat16 rounding against torch's conversion, its Adam step against
torch.optim.Adam. All on the CPU."""
import numpy as np
import pytest
import torch

import mlp_oracle as M
from warpsph_amd import mlp


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 8)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
    return x, t


def test_layout_and_init():
    assert M.N_PARAM == mlp.N_PARAM == 4801
    p = mlp.init_params(3)
    assert p.dtype == np.float32 and p.shape == (4801,)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    assert np.all(b1 == 0) and np.all(b2 == 0) and b3 == 0
    assert np.abs(W1).max() <= np.sqrt(6 / 8) and np.abs(W2).max() <= np.sqrt(6 / 64)


def test_gradient_matches_finite_differences():
    x, t = _data(40, 1)
    p = mlp.init_params(1).astype(np.float64)
    p[M.OFF_B1:M.OFF_W2] = 0.1
    _, g = M.grad(p, x, t, bf16=False)
    rng = np.random.default_rng(0)
    # every bias, plus a sample of each weight block
    idx = np.concatenate([np.arange(M.OFF_B1, M.OFF_W2), np.arange(M.OFF_B2, M.OFF_W3),
                          [M.OFF_B3], rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 40, False),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 80, False),
                          np.arange(M.OFF_W3, M.OFF_B3)])
    eps = 1e-6
    for k in idx:
        q = p.copy(); q[k] += eps
        lp = M.loss(q, x, t, bf16=False)
        q[k] -= 2 * eps
        lm = M.loss(q, x, t, bf16=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(g[k]), (k, fd, g[k])


def test_gradient_matches_torch_autograd_straight_through():
    x, t = _data(64, 2)
    p = mlp.init_params(2)
    lo, g = M.grad(p, x, t, bf16=True)

    def rb(a):  # bf16 rounding with an identity gradient
        return a + (a.to(torch.bfloat16).to(torch.float64) - a).detach()

    P = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1 = P[M.OFF_W1:M.OFF_B1].view(64, 8); b1 = P[M.OFF_B1:M.OFF_W2]
    W2 = P[M.OFF_W2:M.OFF_B2].view(64, 64); b2 = P[M.OFF_B2:M.OFF_W3]
    W3 = P[M.OFF_W3:M.OFF_B3]; b3 = P[M.OFF_B3]
    X = torch.tensor(x, dtype=torch.float64)
    h1 = torch.nn.functional.silu(rb(X) @ rb(W1).T + b1)
    h2 = torch.nn.functional.silu(rb(h1) @ rb(W2).T + b2)
    y = rb(h2) @ rb(W3) + b3
    L = 0.5 * torch.mean((y - torch.tensor(t, dtype=torch.float64)) ** 2)
    L.backward()
    assert abs(L.item() - lo) <= 1e-6 * abs(lo)
    # torch rounds float64 -> bf16 directly, the oracle float64 -> float32 -> bf16:
    # they can differ on rare ties, so compare with a tolerance
    gt = P.grad.numpy()
    assert np.linalg.norm(gt - g) <= 1e-3 * np.linalg.norm(g)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.normal(size=100000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 100000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, -3.0e38], np.float32)])
    a = a.astype(np.float32)
    want = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16(a), want)


def test_adam_matches_torch():
    rng = np.random.default_rng(7)
    p0 = rng.normal(size=50)
    P = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([P], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = rng.normal(size=50)
        P.grad = torch.tensor(g)
        opt.step()
        p, m, v = M.adam(p, m, v, g, step, 1e-3)
        np.testing.assert_allclose(p, P.detach().numpy(), rtol=0, atol=1e-12)


def test_feature_scale_matches_lattice_closed_form(orc):
    """a_ref of the learned start's features (oracle/learn_oracle.py) equals
    the closed form of a_ii for a particle inside a resting lattice at rest
    density, -(dt/d)^2 * 0.42483995 (the same lattice sum pinned in
    test_oracle), and the IISPH oracle's interior a_ii divided by it is
    (rho0/rho)^2."""
    import math
    import learn_oracle as LO
    from helpers import oracle_setup
    from warpsph_amd import config
    for cfg in (config.dambreak_8k(), config.dambreak_1m(), config.torus_4m()):
        want = -(cfg.dt0 / cfg.spacing) ** 2 * 0.42483995
        assert abs(LO.a_ref(cfg, cfg.dt0) / want - 1) < 1e-7
    cfg = config.dambreak_8k().replace(fluid_cells=(12, 12, 12), xsph=0.0, force_iters=2)
    sc, P, wt, gt = oracle_setup(cfg)
    psi = orc.boundary_psi(P, wt, sc.xb)
    _, _, info = orc.step(P, wt, gt, sc.x, np.zeros_like(sc.x), sc.xb, sc.vb, psi)
    k = (6 * 12 + 6) * 12 + 6
    f1 = info["a_ii"][k] / LO.a_ref(cfg, cfg.dt0)
    assert abs(f1 - (cfg.rho0 / info["rho"][k]) ** 2) < 1e-4
    # the device's constant: a_ii * k1 / dt^2 with k1 = -rho0^2 / (m^2 S)
    S = LO.lattice_grad_sq_sum(cfg.h, cfg.spacing)
    k1 = -(cfg.rho0 ** 2) / (cfg.mass ** 2 * S)
    assert abs(k1 * LO.a_ref(cfg, cfg.dt0) / cfg.dt0 ** 2 - 1) < 1e-12
    assert math.isfinite(k1)


def test_start_value_rule():
    import learn_oracle as LO
    from warpsph_amd import config
    cfg = config.dambreak_8k()
    carry = np.array([0.0, 100.0, 100.0, 100.0, 50.0], np.float32)
    y = np.array([0.5, -1.0, np.nan, 0.2, 1.0], np.float32)
    a = np.array([-1e-3, -1e-3, -1e-3, -1e-3, 5e-7], np.float32)
    f1 = np.array([1.0, 1.0, 1.0, 0.01, 1.0], np.float32)
    p = LO.start_value(carry, y, a, cfg, f1)
    P = np.float32(LO.p_ref(cfg))
    assert p[0] == np.float32(0.0) + P * np.float32(0.5)
    assert p[1] == 0.0                      # clamped at 0
    assert p[2] == 100.0                    # non-finite prediction: carried value
    assert p[3] == 100.0                    # weakly coupled: carried value
    assert p[4] == 0.0                      # guard band |a_ii| < inv_small
