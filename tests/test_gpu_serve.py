"""Continuous batching: mc_batch_advance and the request server.

A request's rollout split into chunks, run inside batches that change
membership between chunks, must give bit for bit what the plain rollout of
that request alone gives (state and centre-of-mass displacement).
"""
from __future__ import annotations

import sys
from dataclasses import replace

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


def _act(wl, dev):
    return tuple(torch.as_tensor(a, device=dev) for a in (wl.act_amp, wl.act_omega, wl.act_phase))


def _alone(wl, steps, cfg, dev):
    """The plain single-instance rollout with displacement tracking."""
    from microcosmos_amd import Rollout
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, cfg, wl.dt, actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    tr = ro.track()
    ro.run(steps, use_graph=True)
    torch.cuda.synchronize()
    return ro, _np(tr.cumulative()[0])


def _assert_same(tag, n, e, f, disp, ro, disp_ref):
    np.testing.assert_array_equal(_np(n.position), _np(ro.position), err_msg=f"{tag} pos")
    np.testing.assert_array_equal(_np(n.velocity), _np(ro.velocity), err_msg=f"{tag} vel")
    np.testing.assert_array_equal(_np(e.theta), _np(ro.theta), err_msg=f"{tag} theta")
    np.testing.assert_array_equal(_np(f.steric), _np(ro.steric), err_msg=f"{tag} steric")
    np.testing.assert_array_equal(_np(f.fluid_velocity), _np(ro.fluid_velocity), err_msg=f"{tag} u*")
    np.testing.assert_array_equal(_np(f.f_grid), _np(ro.f_grid), err_msg=f"{tag} f")
    np.testing.assert_array_equal(_np(disp), disp_ref, err_msg=f"{tag} displacement")


def _clone_state(n, e, f):
    return (replace(n, position=n.position.clone(), velocity=n.velocity.clone(),
                    debug_vector=n.debug_vector.clone()),
            replace(e, theta=e.theta.clone()),
            replace(f, steric=f.steric.clone(), fluid_velocity=f.fluid_velocity.clone(),
                    f_grid=f.f_grid.clone()))


def _mix():
    return [W.small(grid=(40, 44), n_fil=3, n_nodes=20, seed=17),
            W.small(grid=(32, 32), seed=5, node_counts=[1, 3, 12, 2]),
            W.small(grid=(48, 48), n_fil=2, n_nodes=16, seed=7),
            W.small(grid=(24, 36), n_fil=1, n_nodes=30, seed=9)]


def test_batch_advance_chunks_and_rejoins(dev):
    """Instances advanced in uneven chunks, the batch re-formed between chunks
    with instances at different clocks, equal their uninterrupted rollouts."""
    _say("mc_batch_advance: chunks 5, 11, 3 with re-formed batches")
    from microcosmos_amd.batch import Batch
    wls = _mix()
    cfg = device_config(wls[0])
    # phase 1: instances 0, 1 for 5 steps; phase 2: 0, 1, 2 (2 joins at 0)
    # for 11 steps; phase 3: 2, 3 (3 joins at 0) for 3 steps
    st = {i: device_state(w, dev) for i, w in enumerate(wls)}
    done = {i: 0 for i in range(4)}
    disp = {i: torch.zeros((1, 2), dtype=torch.float32, device=dev) for i in range(4)}
    for members, k in (([0, 1], 5), ([0, 1, 2], 11), ([2, 3], 3)):
        b = Batch([st[i] for i in members], cfg, wls[0].dt,
                  actuation=[_act(wls[i], dev) for i in members])
        d = torch.cat([disp[i] for i in members]).contiguous()
        b.advance(k, [done[i] for i in members], d)
        torch.cuda.synchronize()
        for j, i in enumerate(members):
            st[i] = _clone_state(*b.instance(j))
            disp[i] = d[j:j + 1].clone()
            done[i] += k
        b.close()
    for i in range(4):
        ro, dref = _alone(wls[i], done[i], cfg, dev)
        n, e, f = st[i]
        _assert_same(f"inst {i} ({done[i]} steps)", n, e, f, disp[i][0], ro, dref)


def _requests(dev, items):
    from microcosmos_amd.serve import Request
    out = []
    for arrival, wl, steps in items:
        nodes, edges, fields = device_state(wl, dev)
        # host topology: plans are built from it without a device read
        edges = replace(edges, pairs=wl.pairs.copy(), bending_pairs=wl.bending_pairs.copy())
        out.append(Request(nodes=nodes, edges=edges, fields=fields, steps=steps,
                           actuation=_act(wl, dev), arrival=arrival))
    return out


def test_server_mix_equals_alone(dev):
    """Requests that arrive while others run, admitted as memory frees up,
    batched with whatever else is running: each result equals the plain
    rollout of that request alone, bit for bit."""
    _say("server: 6 staggered requests, 2 lanes, chunk 16, tight budget")
    from microcosmos_amd.serve import Server, footprint_bytes, latency_report
    wls = _mix() + [W.small(grid=(36, 36), n_fil=2, n_nodes=24, seed=21),
                    W.small(grid=(64, 48), n_fil=4, n_nodes=18, seed=23)]
    steps = [40, 23, 57, 16, 31, 45]
    arrivals = [0.0, 0.0, 0.002, 0.004, 0.004, 0.01]
    cfg = device_config(wls[0])
    fps = [footprint_bytes(w.grid_shape, w.grid_shape, w.num_nodes, w.pairs.shape[0],
                           w.bending_pairs.shape[0]) for w in wls]
    budget = sum(sorted(fps)[-3:])  # at most about three requests resident at once
    srv = Server(cfg, wls[0].dt, hbm_budget=budget, lanes=2, chunk=16, max_batch=3, device=dev)
    reqs = _requests(dev, list(zip(arrivals, wls, steps)))
    done = srv.run(reqs, realtime=True)
    torch.cuda.synchronize()
    assert srv.peak_used <= budget
    assert all(r.finished is not None for r in done)
    rep = latency_report(done)
    _say(f"served: {rep}; reforms {[l.reforms for l in srv.lanes]}")
    assert sum(l.reforms for l in srv.lanes) >= 3
    for r, wl, k in zip(reqs, wls, steps):
        ro, dref = _alone(wl, k, cfg, dev)
        n, e, f = r.result
        _assert_same(f"request {r.rid}", n, e, f, r.displacement, ro, dref)


def test_server_single_request_alone(dev):
    """One request through a one-lane server equals its plain rollout."""
    _say("server: one request, one lane, chunk 7")
    from microcosmos_amd.serve import Server
    wl = W.small(grid=(40, 40), n_fil=2, n_nodes=20, seed=31)
    cfg = device_config(wl)
    srv = Server(cfg, wl.dt, lanes=1, chunk=7, device=dev)
    (r,) = srv.run(_requests(dev, [(0.0, wl, 30)]), realtime=False)
    torch.cuda.synchronize()
    ro, dref = _alone(wl, 30, cfg, dev)
    n, e, f = r.result
    _assert_same("single", n, e, f, r.displacement, ro, dref)


def test_server_gradient_requests_equal_alone(dev):
    """Gradient requests (folding gradients) served next to rollouts give
    the same loss and gradient as computed alone."""
    _say("server: 2 gradient requests among 3 rollouts")
    from microcosmos_amd import PBD_SCHEME_NO_FLUID
    from microcosmos_amd.grad import rollout_value_and_grad
    from microcosmos_amd.serve import Request, Server
    wls = _mix()[:3]
    cfg = device_config(wls[0])
    reqs = _requests(dev, [(0.0, wls[0], 20), (0.001, wls[1], 14), (0.002, wls[2], 9)])
    grads = []
    for k, seed in enumerate((5, 6)):
        wl, tgt = W.folding(seed=seed, grid=(64, 64), n_fil=3, n_nodes=20, steps=60)
        nodes, edges, fields = device_state(wl, dev)
        edges = replace(edges, pairs=wl.pairs.copy(), bending_pairs=wl.bending_pairs.copy())
        grads.append((wl, tgt))
        reqs.append(Request(nodes=nodes, edges=edges, fields=fields, steps=60, arrival=0.0005 * k,
                            gradient=True, target=torch.as_tensor(tgt, device=dev)))
    srv = Server(cfg, 1.0, lanes=2, chunk=8, device=dev, checkpoint_every=16)
    srv.run(reqs, realtime=True)
    torch.cuda.synchronize()
    gcfg = replace(PBD_SCHEME_NO_FLUID, enable_steric=False)
    for r, (wl, tgt) in zip(reqs[3:], grads):
        nodes, edges, fields = device_state(wl, dev)
        loss, g = rollout_value_and_grad(nodes, edges, fields, gcfg, 1.0, 60,
                                         torch.as_tensor(tgt, device=dev), checkpoint_every=16)
        torch.cuda.synchronize()
        assert float(r.loss) == float(loss)
        np.testing.assert_array_equal(_np(r.rest_angle_grad), _np(g))
    for r, wl, k in zip(reqs[:3], wls, (20, 14, 9)):
        ro, dref = _alone(wl, k, cfg, dev)
        n, e, f = r.result
        _assert_same(f"request {r.rid}", n, e, f, r.displacement, ro, dref)
