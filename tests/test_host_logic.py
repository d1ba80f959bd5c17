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


def test_folding_schedule_closed_form():
    """Warmup from lr/100 to lr, then cosine to lr_min (optax linear_schedule
    and cosine_decay_schedule joined at the warmup boundary)."""
    import math
    from microcosmos_amd.grad import folding_schedule
    lr, lo, wu, tot = 0.06, 3e-5, 15, 150
    assert folding_schedule(0, lr, lo, wu, tot) == pytest.approx(lr * 0.01)
    assert folding_schedule(5, lr, lo, wu, tot) == pytest.approx(lr * 0.01 + (lr - lr * 0.01) / 3)
    assert folding_schedule(wu, lr, lo, wu, tot) == pytest.approx(lr)
    mid = wu + (tot - wu) // 2
    t = (mid - wu) / (tot - wu)
    assert folding_schedule(mid, lr, lo, wu, tot) == pytest.approx(
        lr * ((1 - lo / lr) * 0.5 * (1 + math.cos(math.pi * t)) + lo / lr))
    assert folding_schedule(tot, lr, lo, wu, tot) == pytest.approx(lo)
    assert folding_schedule(tot + 50, lr, lo, wu, tot) == pytest.approx(lo)


def test_folding_optimizer_matches_adam_with_clipping():
    """Three updates against a float64 restatement of optax's
    chain(clip_by_global_norm, adam(schedule)): the first gradient is clipped,
    the later ones are not."""
    import torch
    from microcosmos_amd.grad import FoldingOptimizer, folding_schedule
    rng = np.random.default_rng(0)
    p0 = rng.normal(size=12)
    grads = [rng.normal(size=12) * 3.0, rng.normal(size=12) * 0.1, rng.normal(size=12) * 0.05]
    opt = FoldingOptimizer(lr=0.06, lr_min=3e-5, warmup_steps=2, total_steps=10, grad_clip=1.0)
    p = torch.tensor(p0, dtype=torch.float32)
    mu = np.zeros(12)
    nu = np.zeros(12)
    q = p0.copy()
    for k, g in enumerate(grads):
        p = opt.update(p, torch.tensor(g, dtype=torch.float32))
        n = np.sqrt((g * g).sum())
        gc = g if n < 1.0 else g / n
        mu = 0.9 * mu + 0.1 * gc
        nu = 0.999 * nu + 0.001 * gc * gc
        mh = mu / (1 - 0.9 ** (k + 1))
        nh = nu / (1 - 0.999 ** (k + 1))
        q = q - folding_schedule(k, 0.06, 3e-5, 2, 10) * mh / (np.sqrt(nh) + 1e-8)
        np.testing.assert_allclose(p.numpy(), q, rtol=1e-5, atol=1e-6)


def test_folding_optimizer_first_step_is_sign_times_lr():
    """Adam's first bias-corrected step is lr * g / (|g| + eps) per component
    (sign times lr up to eps), whatever the gradient's scale: clipping only
    rescales g (here by its norm, 40.1, so the 1e-3 entry feels eps)."""
    import torch
    from microcosmos_amd.grad import FoldingOptimizer
    g = torch.tensor([3.0, -0.5, 1e-3, -40.0])
    opt = FoldingOptimizer(lr=0.06, warmup_steps=0, total_steps=10)
    p = opt.update(torch.zeros(4), g)
    gc = g.numpy().astype(np.float64) / np.linalg.norm(g.numpy().astype(np.float64))
    np.testing.assert_allclose(p.numpy(), -0.06 * gc / (np.abs(gc) + 1e-8), rtol=1e-5)
    assert np.all(np.abs(np.abs(p.numpy()) - 0.06) < 3e-5)


def test_folding_frame_indices():
    """experiment.py:787-801, 872-877 evaluated by hand at the folding
    configuration (num_steps 150, stability_steps 30, 4 evaluation and 4
    stability frames, anchor offset 30): evaluation 74 + 75 i / 3, stability
    150 + 29 i / 3 rounded, anchor min(149 + 30, 179)."""
    from microcosmos_amd import folding as F
    assert F.eval_indices(150, 4) == [74, 99, 124, 149]
    assert F.eval_indices(150, 1) == [149]
    assert F.eval_indices(9, 3) == [3, 6, 8]
    assert F.stability_indices(150, 30, 4) == [150, 160, 169, 179]
    assert F.stability_indices(150, 30, 1) == [179]
    assert F.stability_indices(150, 0, 4) == []
    assert F.anchor_index(150, 30, 149, 30) == 179
    assert F.anchor_index(150, 10, 149, 30) == 159
    assert F.anchor_index(150, 0, 149, 30) is None
    cfg = F.FoldingLossConfig()
    assert F.objective_steps(150, 30, cfg) == [75, 100, 125, 150, 151, 161, 170, 180]
