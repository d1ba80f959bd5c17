"""Plan autotuning (mc_plan_autotune): the chosen launch knobs change no
result -- a tuned rollout equals an untuned one bit for bit, alone and
batched -- the caller's state is untouched by the timing runs, and a second
plan of the same shape takes the cached choice."""
from __future__ import annotations

import sys

import numpy as np
import pytest

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


def _np(t):
    return t.detach().cpu().numpy()


def _private(ro):
    from microcosmos_amd._runtime import Plan
    ro.plan = Plan(ro.grid_shape, tuple(ro.f_grid.shape[1:]), ro.position.shape[0],
                   ro.plan.pairs, ro.plan.bending_pairs)
    return ro


@pytest.mark.parametrize("case", ["wide", "plain"])
def test_tuned_rollout_equals_untuned(dev, case):
    from microcosmos_amd import Rollout
    if case == "wide":
        wl = W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3)  # one wide cluster
    else:
        wl = W.small(grid=(48, 48), n_fil=3, n_nodes=20, seed=23)
    cfg = device_config(wl)
    act = (wl.act_amp, wl.act_omega, wl.act_phase)
    runs = []
    for tune in (True, False):
        _say(f"{case}: tuned={tune}")
        nodes, edges, fields = device_state(wl, dev)
        ro = _private(Rollout(nodes, edges, fields, cfg, wl.dt, actuation=act))
        if tune:
            before = _np(ro.position).copy(), _np(ro.f_grid).copy()
            choice = ro.autotune(steps=4)
            torch.cuda.synchronize()
            _say(f"  choice {choice}")
            assert set(choice) == {"wide_min", "lbm_pass"}
            np.testing.assert_array_equal(_np(ro.position), before[0])
            np.testing.assert_array_equal(_np(ro.f_grid), before[1])
        ro.run(20)
        torch.cuda.synchronize()
        runs.append((_np(ro.position), _np(ro.velocity), _np(ro.f_grid)))
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)
    # a new plan of the same shape takes the cached choice without timing
    from microcosmos_amd.simulate import tuning
    nodes, edges, fields = device_state(wl, dev)
    ro2 = _private(Rollout(nodes, edges, fields, cfg, wl.dt, actuation=act))
    assert ro2.autotune(steps=4) == choice


def test_tuned_batch_equals_untuned(dev):
    from microcosmos_amd.batch import Batch
    wls = [W.small(grid=(40, 44), n_fil=3, n_nodes=20, seed=17),
           W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3),
           W.small(grid=(48, 48), n_fil=2, n_nodes=16, seed=7)]
    cfg = device_config(wls[0])
    outs = []
    for tune in (True, False):
        _say(f"batch of 3 (one wide cluster), tuned={tune}")
        acts = [tuple(torch.as_tensor(a, device=dev) for a in (w.act_amp, w.act_omega, w.act_phase))
                for w in wls]
        b = Batch([device_state(w, dev) for w in wls], cfg, wls[0].dt, actuation=acts)
        if tune:
            _say(f"  choice {b.autotune(steps=4)}")
        disp = b.run(18)
        torch.cuda.synchronize()
        outs.append((_np(b.position), _np(b.f_grid), _np(disp)))
    for a, c in zip(*outs):
        np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("case", ["plain", "wide"])
def test_lbm_placements_agree(dev, case):
    """The three placements of the LBM pass in a graph-replayed rollout (its
    own launch, inside the next rod launch, on a second stream beside the next
    steric and rod launches) and the step-by-step rollout agree bit for bit."""
    from microcosmos_amd import Rollout
    if case == "wide":
        wl = W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3)
    else:
        wl = W.small(grid=(48, 48), n_fil=3, n_nodes=20, seed=23)
    cfg = device_config(wl)
    act = (wl.act_amp, wl.act_omega, wl.act_phase)
    runs = []
    for mode in (None, 0, 1, 2):
        _say(f"{case}: lbm placement {mode}")
        nodes, edges, fields = device_state(wl, dev)
        ro = _private(Rollout(nodes, edges, fields, cfg, wl.dt, actuation=act))
        if mode is not None:
            ro.set_tuning(lbm_pass=mode)
        ro.run(37, use_graph=mode is not None)
        torch.cuda.synchronize()
        runs.append((_np(ro.position), _np(ro.velocity), _np(ro.f_grid), _np(ro.fluid_velocity)))
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            np.testing.assert_array_equal(a, b)
