"""Batched plans (mc_batch_plan_create / mc_batch_rollout): instances of
different grid, lattice, filament and node counts stepped together must each
give exactly (bit for bit) what they give alone, and match the oracle."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import device_config, device_state, oracle_params, pos_rel_l2, rel_l2, run_oracle

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


def _mix():
    return [W.small(grid=(40, 44), n_fil=3, n_nodes=20, seed=17),
            W.small(grid=(32, 32), seed=5, node_counts=[1, 3, 12, 2]),
            W.small(grid=(48, 48), n_fil=2, n_nodes=16, seed=7),
            W.small(grid=(24, 36), n_fil=1, n_nodes=30, seed=9)]


def _act(wl, dev):
    return tuple(torch.as_tensor(a, device=dev) for a in (wl.act_amp, wl.act_omega, wl.act_phase))


def _np(t):
    return t.detach().cpu().numpy()


def test_batch_equals_instances_alone(dev):
    _say("batch of 4 ragged instances vs each alone, 12 steps")
    from microcosmos_amd import Rollout
    from microcosmos_amd.batch import Batch
    wls = _mix()
    cfg = device_config(wls[0])
    steps = 12
    b = Batch([device_state(w, dev) for w in wls], cfg, wls[0].dt,
              actuation=[_act(w, dev) for w in wls])
    disp = b.run(steps)
    torch.cuda.synchronize()
    for i, w in enumerate(wls):
        nodes, edges, fields = device_state(w, dev)
        ro = Rollout(nodes, edges, fields, cfg, w.dt,
                     actuation=(w.act_amp, w.act_omega, w.act_phase))
        tr = ro.track()
        ro.run(steps, use_graph=True)
        torch.cuda.synchronize()
        bn, be, bf = b.instance(i)
        np.testing.assert_array_equal(_np(bn.position), _np(ro.position), err_msg=f"inst {i} pos")
        np.testing.assert_array_equal(_np(bn.velocity), _np(ro.velocity), err_msg=f"inst {i} vel")
        np.testing.assert_array_equal(_np(be.theta), _np(ro.theta), err_msg=f"inst {i} theta")
        np.testing.assert_array_equal(_np(bf.steric), _np(ro.steric), err_msg=f"inst {i} steric")
        np.testing.assert_array_equal(_np(bf.fluid_velocity), _np(ro.fluid_velocity),
                                      err_msg=f"inst {i} u*")
        np.testing.assert_array_equal(_np(bf.f_grid), _np(ro.f_grid), err_msg=f"inst {i} f")
        np.testing.assert_array_equal(_np(disp[i]), _np(tr.cumulative()[0]),
                                      err_msg=f"inst {i} displacement")


def test_batch_instance_matches_oracle(dev):
    _say("batch instance vs oracle, 10 steps")
    from microcosmos_amd.batch import Batch
    wls = _mix()
    cfg = device_config(wls[0])
    b = Batch([device_state(w, dev) for w in wls], cfg, wls[0].dt,
              actuation=[_act(w, dev) for w in wls])
    b.run(10, track=False)
    torch.cuda.synchronize()
    for i in (0, 3):
        o = run_oracle(wls[i], 10, oracle_params(wls[i]))
        bn, be, bf = b.instance(i)
        assert pos_rel_l2(wls[i].grid_shape, _np(bn.position), o.position) <= 1e-5
        assert rel_l2(_np(bf.f_grid), o.f_grid) <= 1e-5
        assert rel_l2(_np(bf.fluid_velocity), o.fluid_velocity) <= 1e-5
        assert rel_l2(_np(be.theta), o.theta) <= 1e-5


def test_batch_mixed_filter_paths(dev):
    """sigma 1.5: the 64^2 and 24x36 instances take the direct filter (their
    periodic kernels fit the tap limit), the 256^2 instance the FFT; in one
    batch each kernel skips the other path's instances, and every instance
    stays bit-identical to itself alone (steric field included)."""
    _say("batch with direct-filter and FFT instances (sigma 1.5) vs each alone, 6 steps")
    from dataclasses import replace
    from microcosmos_amd import Rollout
    from microcosmos_amd.batch import Batch
    assert R.gauss_taps(256, 1.5) is None and R.gauss_taps(64, 1.5) is not None
    wls = [W.small(grid=(64, 64), n_fil=2, n_nodes=16, seed=3),
           W.small(grid=(256, 256), n_fil=3, n_nodes=24, seed=4),
           W.small(grid=(24, 36), n_fil=1, n_nodes=30, seed=9)]
    cfg = replace(device_config(wls[0]), steric_sigma=1.5)
    steps = 6
    b = Batch([device_state(w, dev) for w in wls], cfg, wls[0].dt,
              actuation=[_act(w, dev) for w in wls])
    b.run(steps, track=False)
    torch.cuda.synchronize()
    for i, w in enumerate(wls):
        nodes, edges, fields = device_state(w, dev)
        ro = Rollout(nodes, edges, fields, cfg, w.dt,
                     actuation=(w.act_amp, w.act_omega, w.act_phase))
        ro.run(steps, use_graph=False)
        torch.cuda.synchronize()
        bn, be, bf = b.instance(i)
        np.testing.assert_array_equal(_np(bf.steric), _np(ro.steric), err_msg=f"inst {i} steric")
        np.testing.assert_array_equal(_np(bn.position), _np(ro.position), err_msg=f"inst {i} pos")
        np.testing.assert_array_equal(_np(bf.f_grid), _np(ro.f_grid), err_msg=f"inst {i} f")
        assert np.abs(_np(bf.steric)).max() > 0.0
