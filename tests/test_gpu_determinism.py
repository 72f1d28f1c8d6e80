"""Run-twice determinism at full size: the 1,000,000-particle dam break with
the learned start (the shipped network), stepped twice from the same inputs,
gives bit-identical positions, velocities, pressures and per-step iteration
counts under the device-driven and the persistent Jacobi drivers. Every
reduction on the path (residual sums, neighbour-count and speed maxima, the
network's MFMA products) is integer or fixed in order, so no result may
depend on the order in which blocks run."""
import sys
from pathlib import Path

import pytest

from helpers import assert_same
from warpsph_amd import config

pytestmark = pytest.mark.gpu

NET = Path(__file__).resolve().parents[1] / "warpsph_amd" / "pressure_net.npy"


@pytest.mark.parametrize("mode", [config.JACOBI_DEVICE, config.JACOBI_PERSISTENT])
def test_run_twice_bit_identical(mode):
    from warpsph_amd.mlp import PressureNet
    from warpsph_amd.solver import IISPH
    cfg = config.dambreak_1m().replace(warm_start=config.WARM_LEARNED, jacobi_mode=mode)
    runs = []
    for rep in range(2):
        print(f"[gpu] run-twice 1M learned start, driver {mode}, run {rep}", flush=True)
        sys.stdout.flush()
        net = PressureNet.load(NET, max_batch=1024)
        s = IISPH(stage_path=None, config=cfg, collect_stats=True, predictor=net)
        its = []
        for _ in range(6):
            s.step()
            its.append(s.step_stats()["iterations"])
        runs.append((its, s.x.cpu().numpy(), s.v.cpu().numpy(), s.p.cpu().numpy()))
        s.close()
    assert runs[0][0] == runs[1][0]
    for name, a, b in zip(("x", "v", "p"), runs[0][1:], runs[1][1:]):
        assert_same(name, a, b)
