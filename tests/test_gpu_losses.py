"""Device folding losses (mc_point_loss) against the float64 restatement
(oracle/losses.py, itself pinned by finite differences on the CPU)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import losses as L

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


def _cloud(seed, n=384, m=600):
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=(n, 2)) * 20.0 + 128.0).astype(np.float32)
    t = (rng.normal(size=(m, 2)) * 15.0 + 120.0).astype(np.float32)
    return p, t


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("name", ["chamfer", "dcd", "smooth_max_nn"])
@pytest.mark.parametrize("size", [(384, 600), (1, 1), (700, 3), (5, 1000)])
def test_loss_matches_restatement(dev, name, size):
    _say(f"{name} {size}")
    from microcosmos_amd import losses as D
    p, t = _cloud(7, *size)
    fn = {"chamfer": D.chamfer_loss, "dcd": D.dcd_loss, "smooth_max_nn": D.smooth_max_nn}[name]
    loss, g = fn(torch.as_tensor(p, device=dev), torch.as_tensor(t, device=dev))
    torch.cuda.synchronize()
    lo, go = getattr(L, name)(p.astype(np.float64), t.astype(np.float64))
    assert abs(float(loss.item()) - lo) <= 1e-5 * abs(lo) + 1e-6
    assert _rel(g.cpu().numpy(), go) <= 1e-5


def test_tied_minimum_splits_the_gradient(dev):
    """A particle exactly between two targets (and a target exactly between
    two particles): the derivative of the tied minimum is split equally."""
    from microcosmos_amd import losses as D
    p = np.array([[0.0, 0.0], [4.0, 0.0], [2.0, 3.0]], np.float32)
    t = np.array([[1.0, 0.0], [-1.0, 0.0], [2.0, 0.0]], np.float32)
    for fn, name in ((D.chamfer_loss, "chamfer"), (D.dcd_loss, "dcd")):
        loss, g = fn(torch.as_tensor(p, device=dev), torch.as_tensor(t, device=dev))
        lo, go = getattr(L, name)(p.astype(np.float64), t.astype(np.float64))
        assert float(loss.item()) == pytest.approx(lo, rel=1e-6)
        np.testing.assert_allclose(g.cpu().numpy(), go, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("name", ["chamfer", "dcd", "smooth_max_nn"])
def test_pca_aligned_matches_restatement(dev, name):
    _say(f"{name} PCA-aligned")
    from microcosmos_amd import losses as D
    p, t = _cloud(9)
    fn = {"chamfer": D.chamfer_loss, "dcd": D.dcd_loss, "smooth_max_nn": D.smooth_max_nn}[name]
    loss, g = D.pca_aligned(fn, torch.as_tensor(p, device=dev), torch.as_tensor(t, device=dev))
    torch.cuda.synchronize()
    lo, go = L.pca_aligned(getattr(L, name), p.astype(np.float64), t.astype(np.float64))
    assert abs(float(loss.item()) - lo) <= 1e-4 * abs(lo)
    assert _rel(g.cpu().numpy(), go) <= 1e-3


def test_chamfer_gradient_drives_the_rod_adjoint(dev):
    """The loss's gradient is the rod adjoint's cotangent: the folding rollout
    (100 PBD steps) with a Chamfer loss against the target curve, against the
    restatement's adjoint along the same float32 trajectory with the float64
    Chamfer restatement as its loss."""
    from dataclasses import replace
    from microcosmos_amd import PBD_SCHEME_NO_FLUID
    from microcosmos_amd import losses as D
    from microcosmos_amd import workloads as W
    from microcosmos_amd.grad import RodAdjoint
    from oracle import rod_adjoint as A
    from oracle import step_ref as R
    from .helpers import device_state, oracle_state
    wl, target = W.folding(steps=100)
    cfg = replace(PBD_SCHEME_NO_FLUID, enable_steric=False)
    nodes, edges, fields = device_state(wl, dev)
    ad = RodAdjoint(nodes, edges, fields, cfg, wl.dt)
    x = ad.forward(100, 25)
    loss, gx = D.chamfer_loss(x, torch.as_tensor(target, device=dev))
    g = ad.backward(gx)[0].cpu().numpy().astype(np.float64)
    s = oracle_state(wl, np.float64)
    p = replace(R.PBD_SCHEME_NO_FLUID, enable_steric=False)
    lo, go = A.rollout_grad(s, wl.dt, p, 100, target, dtype=np.float32,
                            loss_fn=lambda pos: L.chamfer(pos, target.astype(np.float64)))
    rel = _rel(g, go)
    _say(f"chamfer rollout: loss {float(loss.item()):.6f} (restatement {lo:.6f}), rel-L2 {rel:.2e}")
    assert abs(float(loss.item()) - lo) <= 1e-5 * abs(lo)
    assert rel <= 1e-4


@pytest.mark.parametrize("use_anchor", [False, True])
@pytest.mark.parametrize("shape", [(4, 384), (1, 1), (9, 2500)])
def test_frame_stability_matches_restatement(dev, use_anchor, shape):
    """mc_frame_stability against the restatement, with frames that cross
    the periodic boundary (minimum image) on a non-square grid."""
    s, n = shape
    _say(f"frame stability frames={s} n={n} anchor={use_anchor}")
    from microcosmos_amd import losses as D
    rng = np.random.default_rng(s * 1000 + n)
    grid = (192, 256)
    size = np.array([grid[1], grid[0]], np.float64)
    main = (rng.uniform(0, 1, (n, 2)) * size).astype(np.float32)
    fr = np.mod(main[None] + rng.normal(0, 4.0, (s, n, 2)), size).astype(np.float32)
    anchor = np.mod(main + rng.normal(0, 3.0, (n, 2)), size).astype(np.float32) if use_anchor \
        else None
    t = lambda a: None if a is None else torch.as_tensor(a, device=dev)
    l, gf, gm, ga = D.frame_stability(grid, t(fr), t(main), t(anchor))
    torch.cuda.synchronize()
    lo, gfo, gmo, gao = L.frame_stability(grid, fr.astype(np.float64), main.astype(np.float64),
                                          None if anchor is None else anchor.astype(np.float64))
    assert float(l.item()) == pytest.approx(lo, rel=1e-5)
    assert _rel(gf.cpu().numpy(), gfo) <= 1e-5
    assert _rel(gm.cpu().numpy(), gmo) <= 1e-5
    if use_anchor:
        assert _rel(ga.cpu().numpy(), gao) <= 1e-5
    else:
        assert ga is None
