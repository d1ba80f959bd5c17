"""The measurement entry points (mc_profile_step, mc_profile_fused_step)
advance the state exactly as a graph rollout of the same steps does, and
return positive durations; the fused one refuses a scenario without fluid."""
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


def _rollout(wl, dev):
    from microcosmos_amd import Rollout
    nodes, edges, fields = device_state(wl, dev)
    return Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                   actuation=(wl.act_amp, wl.act_omega, wl.act_phase))


@pytest.mark.parametrize("how", ["fused", "unfused"])
def test_profile_matches_rollout(dev, how):
    wl = W.small(grid=(48, 48), n_fil=3, n_nodes=20, seed=23)
    steps = 8
    _say(f"profile {how}: 48x48, 3x20, {steps} steps")
    ref = _rollout(wl, dev)
    ref.run(steps)
    ro = _rollout(wl, dev)
    kern, step = ro.profile_fused(steps) if how == "fused" else ro.profile(steps)
    torch.cuda.synchronize()
    assert kern > 0.0 and step > 0.0 and np.isfinite(kern) and np.isfinite(step)
    assert ro.t == ref.t == steps
    for a, b in ((ro.position, ref.position), (ro.velocity, ref.velocity),
                 (ro.theta, ref.theta), (ro.f_grid, ref.f_grid),
                 (ro.fluid_velocity, ref.fluid_velocity), (ro.steric, ref.steric)):
        np.testing.assert_array_equal(_np(a), _np(b))


def test_profile_fused_needs_fluid_and_even_steps(dev):
    from dataclasses import replace
    from microcosmos_amd._lib import MicrocosmosError
    wl = W.small(grid=(48, 48), n_fil=3, n_nodes=20, seed=23)
    ro = _rollout(wl, dev)
    with pytest.raises(MicrocosmosError, match="num_steps"):
        ro.profile_fused(3)
    nodes, edges, fields = device_state(wl, dev)
    from microcosmos_amd import Rollout
    cfg = replace(device_config(wl), enable_fluid=False)
    dry = Rollout(nodes, edges, fields, cfg, wl.dt)
    with pytest.raises(MicrocosmosError, match="LBM pass"):
        dry.profile_fused(4)
