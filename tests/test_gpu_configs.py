"""The benchmark configurations themselves under the parity bar.

  * the 1024^2 plain case (256 x 64 nodes, 8 PBD cycles): 100 graph-replayed
    steps against the float32 restatement, relative L2 <= 1e-5, with the
    error against the float64 restatement stated;
  * the batched ES population: a seeded slice of the 512-instance population
    (grid sides 128 / 192 / 256 and its most crowded instance) stepped as one
    batch equals each instance alone bit for bit -- state, centre-of-mass
    displacement and its integer binning floor(COM) -- and two instances
    match the restatement at 100 steps;
  * the serving trace: requests of 512^2 and 1024^2 lattices with >= 128
    filaments and a gradient request, served together, each equal to the
    request alone.
"""
from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import (device_config, device_state, oracle_params, oracle_state, pos_rel_l2,
                      rel_l2)

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


@pytest.mark.timeout(900)
def test_plain_1024_100_steps_vs_oracle(dev):
    _say("plain 1024^2 256x64: 100 graph-replayed steps vs the float32 and float64 restatement")
    from microcosmos_amd import Rollout
    wl = W.plain_1024()
    p = oracle_params(wl)
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    ro.run(100)
    torch.cuda.synchronize()
    ro.check()
    s = oracle_state(wl)
    s64 = oracle_state(wl, np.float64)
    for t in range(100):
        s = R.step(replace(s, bending_rest_angles=wl.rest_angles_at(t)), wl.dt, p)
        s64 = R.step(replace(s64, bending_rest_angles=wl.rest_angles_at(t).astype(np.float64)),
                     wl.dt, p)
        if t % 10 == 9:
            _say(f"  oracle step {t + 1}")
    g = {"pos": _np(ro.position), "u": _np(ro.fluid_velocity), "f": _np(ro.f_grid),
         "theta": _np(ro.theta), "steric": _np(ro.steric)}
    r32 = {"pos": pos_rel_l2(wl.grid_shape, g["pos"], s.position),
           "u": rel_l2(g["u"], s.fluid_velocity), "f": rel_l2(g["f"], s.f_grid),
           "theta": rel_l2(g["theta"], s.theta), "steric": rel_l2(g["steric"], s.steric),
           "rho": rel_l2(g["f"].astype(np.float64).sum(0), s.f_grid.astype(np.float64).sum(0))}
    r64 = {"pos": pos_rel_l2(wl.grid_shape, g["pos"], s64.position),
           "u": rel_l2(g["u"], s64.fluid_velocity), "f": rel_l2(g["f"], s64.f_grid),
           "rho": rel_l2(g["f"].astype(np.float64).sum(0), s64.f_grid.sum(0))}
    _say(f"plain1024 100 steps: gpu-vs-fp32 {r32}  gpu-vs-fp64 {r64}")
    for k in ("pos", "u", "f", "theta", "rho"):
        assert r32[k] <= 1e-5, (k, r32)


def _crowded_128():
    """The most crowded instance the population's distribution allows: 16
    filaments of 64 nodes on 128^2 (amplitudes U(0, 0.5), periods
    U(100, 400)); its stencils chain all 1024 nodes into one cluster, which
    the wide IB solve takes."""
    rng = np.random.default_rng(0)
    return W.filaments_workload("es-crowded-128x128-16x64", (128, 128), 16, 0, 1000,
                                node_counts=[64] * 16, amps=rng.uniform(0.0, 0.5, 16),
                                periods=rng.uniform(100.0, 400.0, 16), steps=100,
                                overrides=W._lbm_overrides(0.8, 4))


def _c2_slice():
    """Five instances of each side from the 512-instance population, plus the
    crowded 128^2 instance."""
    pop = W.es_population(512, seed=2, steps=100)
    pick = []
    for side in (128, 192, 256):
        pick += [w for w in pop if w.grid_shape[0] == side][:5]
    crowded = _crowded_128()
    pick.append(crowded)
    return pick, crowded


def _floor_com(wl, pos):
    """Periodic centre of mass (swim_base.py:167-209) and its integer binning
    (utils.py:25-27 index_from_position: floor)."""
    com = R.group_com(pos.astype(np.float32), wl.grid_shape)
    return com, np.floor(com).astype(np.int64)


@pytest.mark.timeout(600)
def test_es_population_slice_batch_equals_alone(dev):
    from microcosmos_amd import Rollout
    from microcosmos_amd.batch import Batch
    wls, crowded = _c2_slice()
    steps = 100
    _say(f"ES population slice: {len(wls)} instances (sides "
         f"{sorted({w.grid_shape[0] for w in wls})}, crowded {crowded.name} with "
         f"{crowded.num_nodes} nodes), {steps} steps batched vs alone")
    cfg = device_config(wls[0])
    b = Batch([device_state(w, dev) for w in wls], cfg, wls[0].dt,
              actuation=[_act(w, dev) for w in wls])
    disp = b.run(steps)
    torch.cuda.synchronize()
    b.check()
    disp = _np(disp)
    for i, w in enumerate(wls):
        nodes, edges, fields = device_state(w, dev)
        ro = Rollout(nodes, edges, fields, cfg, w.dt,
                     actuation=(w.act_amp, w.act_omega, w.act_phase))
        tr = ro.track()
        ro.run(steps)
        torch.cuda.synchronize()
        ro.check()
        bn, be, bf = b.instance(i)
        tag = f"inst {i} {w.name}"
        np.testing.assert_array_equal(_np(bn.position), _np(ro.position), err_msg=f"{tag} pos")
        np.testing.assert_array_equal(_np(bn.velocity), _np(ro.velocity), err_msg=f"{tag} vel")
        np.testing.assert_array_equal(_np(be.theta), _np(ro.theta), err_msg=f"{tag} theta")
        np.testing.assert_array_equal(_np(bf.f_grid), _np(ro.f_grid), err_msg=f"{tag} f")
        np.testing.assert_array_equal(_np(bf.fluid_velocity), _np(ro.fluid_velocity),
                                      err_msg=f"{tag} u*")
        np.testing.assert_array_equal(disp[i], _np(tr.cumulative()[0]), err_msg=f"{tag} disp")
        c_b, bin_b = _floor_com(w, _np(bn.position))
        c_a, bin_a = _floor_com(w, _np(ro.position))
        np.testing.assert_array_equal(bin_b, bin_a, err_msg=f"{tag} floor(COM)")
        assert np.isfinite(disp[i]).all()
        del ro


@pytest.mark.timeout(600)
def test_es_population_slice_matches_oracle(dev):
    """Two population instances (a 192^2 and the crowded 128^2 one) at 100
    steps against the float32 restatement: fitness (cumulative displacement)
    within relative 1e-5, its floor binning exactly, the state within
    relative L2 1e-5."""
    from microcosmos_amd.batch import Batch
    wls, crowded = _c2_slice()
    pick = [next(w for w in wls if w.grid_shape[0] == 192), crowded]
    steps = 100
    _say(f"ES population: {[w.name for w in pick]} batched, {steps} steps vs the restatement")
    cfg = device_config(pick[0])
    b = Batch([device_state(w, dev) for w in pick], cfg, pick[0].dt,
              actuation=[_act(w, dev) for w in pick])
    disp = _np(b.run(steps))
    torch.cuda.synchronize()
    b.check()
    for i, w in enumerate(pick):
        p = oracle_params(w)
        s = oracle_state(w)
        pos_hist = []
        for t in range(steps):
            s = R.step(replace(s, bending_rest_angles=w.rest_angles_at(t)), w.dt, p)
            pos_hist.append(s.position.copy())
        hist = R.swim_track(np.concatenate([oracle_state(w).position[None], np.stack(pos_hist)]),
                            w.grid_shape)
        cum = hist[-1][0]
        bn, be, bf = b.instance(i)
        r = {"pos": pos_rel_l2(w.grid_shape, _np(bn.position), s.position),
             "f": rel_l2(_np(bf.f_grid), s.f_grid), "u": rel_l2(_np(bf.fluid_velocity),
                                                                s.fluid_velocity),
             "disp": rel_l2(disp[i], cum)}
        _say(f"  {w.name}: {r}")
        for k in ("pos", "f", "u", "disp"):
            assert r[k] <= 1e-5, (w.name, k, r)
        _, bin_g = _floor_com(w, _np(bn.position))
        _, bin_o = _floor_com(w, s.position)
        np.testing.assert_array_equal(bin_g, bin_o)


def _served_items():
    """Two requests of each of the trace's two largest sides (512^2, 1024^2),
    each with 128-256 filaments of 16-64 nodes (the trace's distribution,
    seeded), shortened to 120-200 steps."""
    rng = np.random.default_rng(44)
    out = []
    for k, side in enumerate((512, 1024, 512, 1024)):
        nf = int(rng.integers(128, 257))
        wl = W.filaments_workload(f"req-{k}-{side}x{side}-{nf}f", (side, side), nf, 0,
                                  int(rng.integers(1 << 30)),
                                  node_counts=rng.integers(16, 65, size=nf).tolist(),
                                  amps=rng.uniform(0.0, 0.5, nf),
                                  periods=rng.uniform(100.0, 400.0, nf), steps=200,
                                  overrides=W._lbm_overrides(0.8, 4))
        out.append(W.ServeItem(arrival=0.001 * k, workload=wl, steps=int(rng.integers(120, 201)),
                               gradient=False))
    return out


@pytest.mark.timeout(600)
def test_served_large_requests_equal_alone(dev):
    from microcosmos_amd import PBD_SCHEME_NO_FLUID, Rollout
    from microcosmos_amd.grad import rollout_value_and_grad
    from microcosmos_amd.serve import Request, Server
    items = _served_items()
    cfg = device_config(items[0].workload)
    reqs = []
    for k, it in enumerate(items):
        wl = it.workload
        nodes, edges, fields = device_state(wl, dev)
        edges = replace(edges, pairs=wl.pairs.copy(), bending_pairs=wl.bending_pairs.copy())
        reqs.append(Request(nodes=nodes, edges=edges, fields=fields, steps=it.steps,
                            actuation=_act(wl, dev), arrival=0.001 * k))
    gwl, gtgt = W.folding(seed=11, steps=100)
    gn, ge, gf = device_state(gwl, dev)
    ge = replace(ge, pairs=gwl.pairs.copy(), bending_pairs=gwl.bending_pairs.copy())
    reqs.append(Request(nodes=gn, edges=ge, fields=gf, steps=100, arrival=0.0015, gradient=True,
                        target=torch.as_tensor(gtgt, device=dev)))
    _say("served: " + ", ".join(f"{it.workload.name} ({it.workload.num_nodes} nodes, "
                                f"{it.steps} steps)" for it in items) + " + 1 gradient request")
    srv = Server(cfg, items[0].workload.dt, lanes=2, chunk=32, max_batch=4, device=dev,
                 checkpoint_every=25)
    srv.run(reqs, realtime=True)
    torch.cuda.synchronize()
    for r, it in zip(reqs, items):
        wl = it.workload
        nodes, edges, fields = device_state(wl, dev)
        ro = Rollout(nodes, edges, fields, cfg, wl.dt,
                     actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
        tr = ro.track()
        ro.run(it.steps)
        torch.cuda.synchronize()
        ro.check()
        n, e, f = r.result
        tag = f"request {r.rid} {wl.name}"
        np.testing.assert_array_equal(_np(n.position), _np(ro.position), err_msg=f"{tag} pos")
        np.testing.assert_array_equal(_np(n.velocity), _np(ro.velocity), err_msg=f"{tag} vel")
        np.testing.assert_array_equal(_np(f.f_grid), _np(ro.f_grid), err_msg=f"{tag} f")
        np.testing.assert_array_equal(_np(r.displacement), _np(tr.cumulative()[0]),
                                      err_msg=f"{tag} disp")
        _say(f"  {tag}: equal to alone")
        del ro
    gcfg = replace(PBD_SCHEME_NO_FLUID, enable_steric=False)
    loss, g = rollout_value_and_grad(*device_state(gwl, dev), gcfg, 1.0, 100,
                                     torch.as_tensor(gtgt, device=dev), checkpoint_every=25)
    torch.cuda.synchronize()
    assert float(reqs[-1].loss) == float(loss)
    np.testing.assert_array_equal(_np(reqs[-1].rest_angle_grad), _np(g))
