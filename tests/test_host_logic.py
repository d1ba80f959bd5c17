"""Host-side logic of the serving, gradient and workload modules (CPU only,
no device calls)."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import step_ref as R
from microcosmos_amd import workloads as W
from microcosmos_amd.grad import hinge_center_nodes, min_image, node_gradient
from microcosmos_amd.serve import Request, fifo_report, footprint_bytes, latency_report

ROOT = Path(__file__).resolve().parents[1]


def test_fifo_report_closed_form():
    # arrivals 0, 1, 1.5; service 2, 1, 3: finishes 2, 3, 6
    rep = fifo_report(np.array([0.0, 1.0, 1.5]), np.array([2.0, 1.0, 3.0]))
    assert rep["makespan_s"] == pytest.approx(6.0)
    assert rep["rollouts_per_s"] == pytest.approx(0.5)
    # latencies 2, 2, 4.5
    assert rep["p50_latency_s"] == pytest.approx(2.0)
    assert rep["p99_latency_s"] == pytest.approx(np.percentile([2.0, 2.0, 4.5], 99))


def test_fifo_report_idle_gap():
    # the server idles between requests that arrive after the previous finished
    rep = fifo_report(np.array([0.0, 10.0]), np.array([1.0, 1.0]))
    assert rep["p99_latency_s"] == pytest.approx(1.0)
    assert rep["makespan_s"] == pytest.approx(11.0)


def test_latency_report():
    reqs = []
    for a, adm, fin in ((0.0, 0.1, 1.0), (0.5, 0.5, 2.5), (1.0, 2.0, 3.0)):
        r = Request(nodes=None, edges=None, fields=None, steps=1, arrival=a)
        r.admitted, r.finished = adm, fin
        reqs.append(r)
    rep = latency_report(reqs)
    assert rep["p50_latency_s"] == pytest.approx(2.0)
    assert rep["mean_queue_wait_s"] == pytest.approx((0.1 + 0.0 + 1.0) / 3)
    assert rep["rollouts_per_s"] == pytest.approx(1.0)


def test_footprint_grows_with_every_dimension():
    base = footprint_bytes((128, 128), (128, 128), 100, 99, 98)
    assert footprint_bytes((256, 256), (256, 256), 100, 99, 98) > base
    assert footprint_bytes((128, 128), (128, 128), 1000, 999, 998) > base
    # at least the state and the ping-pong lattice
    assert base >= 128 * 128 * (36 + 36 + 8)


def test_serving_trace_is_seeded_and_mixed():
    a = W.serving_trace(24, seed=4)
    b = W.serving_trace(24, seed=4)
    assert [x.arrival for x in a] == [x.arrival for x in b]
    assert all(np.array_equal(x.workload.position, y.workload.position) for x, y in zip(a, b))
    arr = np.array([x.arrival for x in a])
    assert np.all(np.diff(arr) > 0)
    grads = [x for x in a if x.gradient]
    rolls = [x for x in a if not x.gradient]
    assert grads and rolls
    for x in grads:
        assert x.target.shape == x.workload.position.shape and x.steps == 500
    for x in rolls:
        assert x.workload.grid_shape[0] in (128, 256, 512, 1024)
        assert 200 <= x.steps <= 2000


def test_folding_workload_target():
    wl, tgt = W.folding()
    assert wl.num_nodes == 384 and wl.bending_pairs.shape[0] == 368
    d = R.displacement(wl.grid_shape, tgt.astype(np.float64), wl.position.astype(np.float64))
    # each filament is displaced along its normal by at most the amplitude
    assert np.max(np.linalg.norm(d, axis=1)) <= 5.0 + 1e-4
    assert np.all((tgt >= 0) & (tgt <= 256))


def test_hinge_centers_and_node_gradient():
    wl, _ = W.folding(n_fil=2, n_nodes=6)
    c = hinge_center_nodes(wl.pairs, wl.bending_pairs)
    np.testing.assert_array_equal(c, wl.pairs[wl.bending_pairs[:, 0], 1])
    g = torch.arange(c.shape[0], dtype=torch.float32)
    gn = node_gradient(g, c, wl.num_nodes)
    expect = np.zeros(wl.num_nodes, np.float32)
    np.add.at(expect, c, np.arange(c.shape[0], dtype=np.float32))
    np.testing.assert_array_equal(gn.numpy(), expect)
    # filament end nodes carry no hinge
    assert gn[0] == 0 and gn[5] == 0 and gn[6] == 0 and gn[11] == 0


def test_min_image_matches_restatement():
    rng = np.random.default_rng(0)
    a = rng.uniform(0, 64, (50, 2)).astype(np.float32)
    b = rng.uniform(0, 64, (50, 2)).astype(np.float32)
    got = min_image((48, 64), torch.as_tensor(a), torch.as_tensor(b)).numpy()
    np.testing.assert_allclose(got, R.displacement((48, 64), a, b), atol=1e-5)


def test_bench_rejects_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
