"""simulate() on device-recorded, graph-replayed rollouts (mc_rollout_record,
simulate.py:15-35): the frames equal the step-by-step recording bit for bit,
a strided recording keeps every k-th frame, and against the oracle the
per-step positions and the grid-resolution fluid velocity (linearly resized
from a lattice of another resolution, simulate.py:24-28) agree to 1e-5."""
from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import device_config, device_state, oracle_params, oracle_state, pos_rel_l2, rel_l2

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


def _np(t):
    return t.detach().cpu().numpy()


def _with_lattice(wl, hf, wf, seed=3):
    """The workload on a lattice of hf x wf cells (the grid unchanged)."""
    rng = np.random.default_rng(seed)
    u0 = rng.uniform(-1e-3, 1e-3, size=(2, hf, wf)).astype(np.float32)
    return replace(wl, f_grid=W.feq_init(u0), fluid_velocity=np.zeros((2, hf, wf), np.float32))


def test_recorded_frames_equal_step_by_step_recording(dev):
    from microcosmos_amd import Rollout
    wl = W.small(grid=(48, 48), n_fil=3, n_nodes=20, seed=23)
    cfg = device_config(wl)
    steps = 37
    _say(f"graph-recorded vs step-by-step recorded rollout, {steps} steps")
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, cfg, wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    fr = ro.record(steps)
    torch.cuda.synchronize()
    nodes, edges, fields = device_state(wl, dev)
    ro2 = Rollout(nodes, edges, fields, cfg, wl.dt,
                  actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    rp = torch.empty((steps, wl.num_nodes, 2), dtype=torch.float32, device=dev)
    rv = torch.empty((steps, 2, 48, 48), dtype=torch.float32, device=dev)
    ro2.run(steps, use_graph=False, record_position=rp, record_fluid_velocity=rv)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(fr["position"]), _np(rp))
    np.testing.assert_array_equal(_np(fr["fluid_velocity"]), _np(rv))
    np.testing.assert_array_equal(_np(fr["position"][-1]), _np(ro.position))
    np.testing.assert_array_equal(_np(fr["velocity"][-1]), _np(ro.velocity))
    np.testing.assert_array_equal(_np(fr["steric"][-1]), _np(ro.steric))
    np.testing.assert_array_equal(_np(fr["debug_vector"][-1]), _np(ro.debug_vector))
    np.testing.assert_array_equal(_np(ro.position), _np(ro2.position))
    # a strided recording of the same rollout keeps every 4th frame
    nodes, edges, fields = device_state(wl, dev)
    ro3 = Rollout(nodes, edges, fields, cfg, wl.dt,
                  actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    f4 = ro3.record(steps, stride=4)
    torch.cuda.synchronize()
    assert f4["position"].shape[0] == steps // 4
    np.testing.assert_array_equal(_np(f4["position"]), _np(fr["position"][3::4][:steps // 4]))
    np.testing.assert_array_equal(_np(f4["steric"]), _np(fr["steric"][3::4][:steps // 4]))


@pytest.mark.parametrize("lattice", [(48, 48), (96, 96), (32, 40)])
def test_simulate_matches_oracle(dev, lattice):
    from microcosmos_amd import simulate
    wl = W.small(grid=(48, 48), n_fil=2, n_nodes=16, seed=7)
    if lattice != (48, 48):
        wl = _with_lattice(wl, *lattice)
    cfg = device_config(wl)
    steps = 12
    _say(f"simulate() with lattice {lattice} on a 48x48 grid, {steps} steps, vs the oracle")
    nodes, edges, fields = device_state(wl, dev)
    nts, fts = simulate(nodes, edges, fields, wl.dt, steps, cfg)
    torch.cuda.synchronize()
    assert fts.f_grid is None
    assert tuple(fts.fluid_velocity.shape) == (steps, 2, 48, 48)
    assert tuple(nts.position.shape) == (steps, wl.num_nodes, 2)
    assert tuple(nts.color_bend.shape) == (steps, wl.num_nodes)
    s = oracle_state(wl)
    p = oracle_params(wl)
    for k in range(steps):
        s = R.step(s, wl.dt, p)
        assert pos_rel_l2(wl.grid_shape, _np(nts.position[k]), s.position) <= 1e-5, k
        want = R.resize_linear(s.fluid_velocity, (2, 48, 48))
        assert rel_l2(_np(fts.fluid_velocity[k]), want) <= 1e-5, k
        assert rel_l2(_np(fts.steric[k]), s.steric) <= 1e-5, k
