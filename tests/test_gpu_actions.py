"""Per-step actions (mc_rollout_actions, gym/base.py:163-176): before step k
the rest lengths and rest angles are the base values plus action row k.  The
graph-replayed rollout must equal a loop of single steps (mc_step) on edges
rewritten by the host each step, bit for bit, alone and batched, and match
the oracle stepping the same rewritten edges."""
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


def _actions(wl, steps, seed, rest=True, ang=True):
    rng = np.random.default_rng(seed)
    E, B = wl.pairs.shape[0], wl.bending_pairs.shape[0]
    dl = (rng.normal(size=(steps, E)) * 0.05).astype(np.float32) if rest else None
    da = (rng.normal(size=(steps, B)) * 0.2).astype(np.float32) if ang else None
    return dl, da


def _step_loop(wl, cfg, dl, da, steps, dev):
    """The reference's env loop: edges = base + action row, then one step."""
    from microcosmos_amd import step
    nodes, edges, fields = device_state(wl, dev)
    base_l, base_a = edges.rest_lengths.clone(), edges.bending_rest_angles.clone()
    for k in range(steps):
        e = replace(edges,
                    rest_lengths=base_l + torch.as_tensor(dl[k], device=dev) if dl is not None
                    else base_l,
                    bending_rest_angles=base_a + torch.as_tensor(da[k], device=dev)
                    if da is not None else base_a)
        nodes, edges, fields = step(nodes, e, fields, wl.dt, cfg)
    torch.cuda.synchronize()
    return nodes, edges, fields


@pytest.mark.parametrize("which", ["both", "rest_only", "angle_only"])
def test_rollout_actions_equal_step_loop(dev, which):
    from microcosmos_amd import Rollout
    wl = W.small(grid=(48, 48), n_fil=3, n_nodes=20, seed=23)
    cfg = device_config(wl)
    steps = 37  # two 16-step graph replays and a 5-step remainder
    dl, da = _actions(wl, steps, 1, rest=which != "angle_only", ang=which != "rest_only")
    _say(f"actions {which}: rollout of {steps} rows vs loop of mc_step")
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, cfg, wl.dt)
    ro.run_actions(None if dl is None else torch.as_tensor(dl, device=dev),
                   None if da is None else torch.as_tensor(da, device=dev))
    torch.cuda.synchronize()
    n2, e2, f2 = _step_loop(wl, cfg, dl, da, steps, dev)
    np.testing.assert_array_equal(_np(ro.position), _np(n2.position))
    np.testing.assert_array_equal(_np(ro.velocity), _np(n2.velocity))
    np.testing.assert_array_equal(_np(ro.theta), _np(e2.theta))
    np.testing.assert_array_equal(_np(ro.f_grid), _np(f2.f_grid))
    np.testing.assert_array_equal(_np(ro.fluid_velocity), _np(f2.fluid_velocity))
    # the base values are not modified
    np.testing.assert_array_equal(_np(edges.rest_lengths), wl.rest_lengths)


def test_batch_actions_equal_instances_alone_and_oracle(dev):
    from microcosmos_amd import Rollout
    from microcosmos_amd.batch import Batch
    wls = [W.small(grid=(40, 44), n_fil=3, n_nodes=20, seed=17),
           W.small(grid=(32, 32), seed=5, node_counts=[1, 3, 12, 2]),
           W.small(grid=(48, 48), n_fil=2, n_nodes=16, seed=7)]
    cfg = device_config(wls[0])
    steps = 20
    acts = [_actions(w, steps, 10 + i) for i, w in enumerate(wls)]
    dl = np.concatenate([a[0] for a in acts], axis=1)
    da = np.concatenate([a[1] for a in acts], axis=1)
    _say(f"batch of {len(wls)} with per-step actions, {steps} steps, vs each alone")
    b = Batch([device_state(w, dev) for w in wls], cfg, wls[0].dt)
    disp = b.run_actions(torch.as_tensor(dl, device=dev), torch.as_tensor(da, device=dev))
    torch.cuda.synchronize()
    assert torch.isfinite(disp).all()
    for i, w in enumerate(wls):
        nodes, edges, fields = device_state(w, dev)
        ro = Rollout(nodes, edges, fields, cfg, w.dt)
        ro.run_actions(torch.as_tensor(acts[i][0], device=dev),
                       torch.as_tensor(acts[i][1], device=dev))
        torch.cuda.synchronize()
        bn, be, bf = b.instance(i)
        np.testing.assert_array_equal(_np(bn.position), _np(ro.position), err_msg=f"inst {i}")
        np.testing.assert_array_equal(_np(be.theta), _np(ro.theta), err_msg=f"inst {i}")
        np.testing.assert_array_equal(_np(bf.f_grid), _np(ro.f_grid), err_msg=f"inst {i}")
    # instance 0 against the oracle stepping the rewritten edges
    w = wls[0]
    s = oracle_state(w)
    p = oracle_params(w)
    for k in range(steps):
        s = replace(s, rest_lengths=w.rest_lengths + acts[0][0][k],
                    bending_rest_angles=w.bending_rest_angles + acts[0][1][k])
        s = R.step(s, w.dt, p)
    bn, be, bf = b.instance(0)
    assert pos_rel_l2(w.grid_shape, _np(bn.position), s.position) <= 1e-5
    assert rel_l2(_np(be.theta), s.theta) <= 1e-5
    assert rel_l2(_np(bf.f_grid), s.f_grid) <= 1e-5


def test_action_rows_are_checked(dev):
    from microcosmos_amd import Rollout
    wl = W.small(grid=(32, 32), n_fil=1, n_nodes=10, seed=3)
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt)
    with pytest.raises(ValueError):
        ro.run_actions(torch.zeros((4, wl.pairs.shape[0] + 1), device=dev))
    with pytest.raises(ValueError):
        ro.run_actions(None, None)
