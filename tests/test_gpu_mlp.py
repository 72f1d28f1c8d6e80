"""BF16 matrix-core pressure predictor (csrc/mlp.hip) against the float64
restatement oracle/mlp_oracle.py: forward values, loss, gradient, Adam step.

Tolerances: the device rounds the same operands to bf16 as the oracle but
accumulates in float32 and evaluates exp with the hardware approximation, so
a hidden activation can round to the neighbouring bf16 value now and then.
Forward |dy| <= 2e-3 (1 + |y|); loss within 1e-3 relative; the gradient also
rounds dz and h1 to bf16 inside the backward products, so it is compared as a
vector: ||g_gpu - g_cpu|| <= 2e-2 ||g_cpu|| and every block's cosine > 0.999.
"""
import sys

import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 8)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 7]).astype(np.float32)
    return x, t


@pytest.fixture(scope="module")
def torch():
    import torch as T
    return T


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 70001])
def test_forward_matches_oracle(torch, n):
    from warpsph_amd.mlp import PressureNet, init_params
    _announce(f"mlp forward n={n}")
    p = init_params(11)
    rng = np.random.default_rng(1)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(size=64) * 0.1
    p[M.OFF_B2:M.OFF_W3] = rng.normal(size=64) * 0.1
    p[M.OFF_B3] = 0.25
    net = PressureNet(max_batch=1024, params=p)
    x, _ = _data(n, n)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    want = M.forward(p, x)
    assert np.all(np.isfinite(y))
    err = np.abs(y - want)
    assert np.all(err <= 2e-3 * (1 + np.abs(want))), (err.max(), np.argmax(err))
    # typical error is far below the bound
    assert np.median(err) <= 2e-5 * (1 + np.median(np.abs(want)))
    net.close()


@pytest.mark.parametrize("n", [7, 64, 4097, 65536])
def test_loss_and_gradient_match_oracle(torch, n):
    from warpsph_amd.mlp import PressureNet, init_params
    _announce(f"mlp gradient n={n}")
    p = init_params(5)
    net = PressureNet(max_batch=65536, params=p)
    x, t = _data(n, 100 + n)
    loss, g = net.train_step(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), lr=0.0,
                             want_grad=True)
    lo, go = M.grad(p, x, t)
    assert abs(loss - lo) <= 1e-3 * lo
    assert np.linalg.norm(g - go) <= 2e-2 * np.linalg.norm(go)
    blocks = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
              (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAM)]
    for a, b in blocks:
        u, v = g[a:b].astype(np.float64), go[a:b]
        if b - a == 1:
            assert abs(u[0] - v[0]) <= 1e-2 * abs(v[0]) + 1e-6
        else:
            cos = u @ v / (np.linalg.norm(u) * np.linalg.norm(v))
            assert cos > 0.999, (a, b, cos)
    # lr = 0 leaves the parameters alone
    np.testing.assert_array_equal(net.params(), p)
    net.close()


def test_adam_steps_match_oracle(torch):
    from warpsph_amd.mlp import PressureNet, init_params
    _announce("mlp adam")
    p = init_params(9)
    net = PressureNet(max_batch=4096, params=p)
    x, t = _data(4096, 9)
    X, T = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    q = p.astype(np.float64)
    m = np.zeros(M.N_PARAM)
    v = np.zeros(M.N_PARAM)
    for step in range(1, 4):
        _, g = net.train_step(X, T, lr=1e-3, want_grad=True)
        q, m, v = M.adam(q, m, v, g.astype(np.float64), step, 1e-3)
        got = net.params()
        # float32 Adam arithmetic against float64 on the same gradient
        assert np.abs(got - q).max() <= 1e-6, np.abs(got - q).max()
    net.close()


def test_training_reduces_loss(torch):
    from warpsph_amd.mlp import PressureNet
    _announce("mlp training")
    net = PressureNet(max_batch=8192)
    x, t = _data(8192, 3)
    X, T = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    first = net.train_step(X, T, lr=0.0)
    for _ in range(400):
        last = net.train_step(X, T, lr=3e-3)
    assert last < 0.1 * first, (first, last)
    # the trained parameters still agree with the oracle's forward pass
    y = net.forward(X).cpu().numpy()
    want = M.forward(net.params(), x)
    assert np.abs(y - want).max() <= 2e-3 * (1 + np.abs(want).max())
    net.close()
