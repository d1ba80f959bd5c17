"""Swimmer fitness on the device (mc_center_of_mass, the per-step tracker of
mc_track_*) against the oracle's restatement: bit-exact (same summation
order, same floor-mod)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import device_config, device_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


@pytest.mark.parametrize("periodic", [True, False])
def test_center_of_mass_groups(dev, periodic):
    _say(f"center of mass periodic={periodic}")
    from microcosmos_amd.fitness import center_of_mass
    rng = np.random.default_rng(11)
    grid = (64, 96)
    pos = rng.uniform(0, 1, (1500, 2)) * np.array([96.0, 64.0])
    # group 2 is a blob straddling both boundaries
    pos[3:300] = np.mod(rng.uniform(-4, 4, (297, 2)) + np.array([95.0, 63.0]),
                        np.array([96.0, 64.0]))
    pos = pos.astype(np.float32)
    off = [0, 3, 300, 301, 1500]
    got = center_of_mass(torch.as_tensor(pos, device=dev), grid, off, periodic)
    exp = R.group_com(pos, grid, off, periodic)
    np.testing.assert_array_equal(got.cpu().numpy(), exp)


def test_rollout_tracker(dev):
    _say("rollout displacement tracker, record and graph paths")
    from microcosmos_amd import Rollout
    wl = W.plain_128()
    cfg = device_config(wl)
    steps = 24

    def make():
        nodes, edges, fields = device_state(wl, dev)
        return Rollout(nodes, edges, fields, cfg, wl.dt,
                       actuation=(wl.act_amp, wl.act_omega, wl.act_phase))

    # recorded trajectory (step-by-step path) with tracking
    ra = make()
    pos0 = ra.position.clone()
    tr_a = ra.track(start_step=5)
    rec = torch.empty((steps, ra.position.shape[0], 2), dtype=torch.float32, device=dev)
    ra.run(steps, record_position=rec)
    # graph-replayed rollout with tracking
    rb = make()
    tr_b = rb.track(start_step=5)
    rb.run(steps, use_graph=True)
    torch.cuda.synchronize()
    frames = [pos0.cpu().numpy()] + list(rec.cpu().numpy())
    hist = R.swim_track(frames, wl.grid_shape, start_step=5)
    np.testing.assert_array_equal(tr_a.cumulative().cpu().numpy(), hist[-1])
    np.testing.assert_array_equal(tr_a.com().cpu().numpy(),
                                  R.group_com(frames[-1], wl.grid_shape))
    np.testing.assert_array_equal(rb.position.cpu().numpy(), frames[-1])
    np.testing.assert_array_equal(tr_b.cumulative().cpu().numpy(), hist[-1])
    assert abs(tr_a.displacement() - float(np.linalg.norm(hist[-1][0]))) < 1e-6
