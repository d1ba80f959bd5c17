"""The folding-loss restatement (oracle/losses.py) against central finite
differences and closed forms.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import losses as L


def _cloud(seed, n=40, m=55):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 2)) * 5.0 + np.array([3.0, -1.0])
    t = rng.normal(size=(m, 2)) * 4.0
    return p, t


def _fd(f, p, h=1e-6):
    g = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e.flat[i] = h
        g.flat[i] = (f(p + e) - f(p - e)) / (2 * h)
    return g


@pytest.mark.parametrize("name", ["chamfer", "dcd", "smooth_max_nn"])
def test_loss_gradient_matches_fd(name):
    p, t = _cloud(1)
    fn = getattr(L, name)
    loss, g = fn(p, t)
    fd = _fd(lambda x: fn(x, t)[0], p)
    assert np.isfinite(loss)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd)))


@pytest.mark.parametrize("name", ["chamfer", "dcd", "smooth_max_nn"])
def test_pca_aligned_gradient_matches_fd(name):
    """The angle is held fixed (stop_gradient), so the finite differences
    use the angle of the unperturbed cloud."""
    p, t = _cloud(2)
    fn = getattr(L, name)
    a = L.pca_rotation_angle(p, t)
    loss, g = L.pca_aligned(fn, p, t, angle=a)
    fd = _fd(lambda x: L.pca_aligned(fn, x, t, angle=a)[0], p)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd)))


def test_chamfer_closed_forms():
    # identical clouds: zero loss; one point against one point: 2 |p - t|^2
    p, _ = _cloud(3)
    assert L.chamfer(p, p)[0] == 0.0
    l, g = L.chamfer(np.array([[1.0, 2.0]]), np.array([[4.0, 6.0]]))
    assert l == pytest.approx(2 * 25.0)
    np.testing.assert_allclose(g, [[2 * 2 * -3.0, 2 * 2 * -4.0]])


def test_dcd_counts_divide_shared_targets():
    """Three particles all nearest to one of two targets: that target's
    count is 3, so the particle terms are d / 3; brute force."""
    p = np.array([[0.0, 0.0], [0.1, 0.0], [0.0, 0.2]])
    t = np.array([[0.05, 0.05], [10.0, 10.0]])
    d2 = ((p[:, None] - t[None]) ** 2).sum(-1)
    exp_pt = (d2[:, 0] / 3.0).mean()
    # targets -> particles: target 0 nearest particle 0 or 1 (tie-free here), target 1 nearest 2
    nn_t = d2.argmin(axis=0)
    cnt = np.bincount(nn_t, minlength=3).astype(float)
    exp_tp = (d2.min(axis=0) / cnt[nn_t]).mean()
    assert L.dcd(p, t)[0] == pytest.approx(exp_pt + exp_tp)


def test_min_tie_splits_the_gradient():
    """A particle equidistant from two targets: jax's min rule splits the
    gradient equally, so the particle-side term pulls towards their midpoint."""
    p = np.array([[0.0, 0.0]])
    t = np.array([[1.0, 0.0], [-1.0, 0.0]])
    l, g = L.chamfer(p, t)
    # particle term: 1 * 2 (p - midpoint) = 0; target terms: each target's
    # only particle is p: (2 / 2) (p - t_j) summed = -(t_0 + t_1) = 0
    np.testing.assert_allclose(g, [[0.0, 0.0]], atol=1e-15)
    assert l == pytest.approx(1.0 + 1.0)


def test_smooth_max_approaches_max():
    p, t = _cloud(4)
    d = np.sqrt(((p[:, None] - t[None]) ** 2).sum(-1).min(axis=1) + 1e-8)
    for a in (1.0, 10.0, 100.0):
        l, _ = L.smooth_max_nn(p, t, alpha=a)
        assert d.max() <= l <= d.max() + np.log(p.shape[0]) / a + 1e-12


def test_pca_angle_of_rotated_copy():
    """A cloud and its copy rotated by 0.7 rad: the principal axes differ by
    0.7 (mod pi), and the aligned chamfer loss is ~0."""
    rng = np.random.default_rng(5)
    t = rng.normal(size=(80, 2)) * np.array([6.0, 1.0])
    c, s = np.cos(-0.7), np.sin(-0.7)
    p = (t - t.mean(0)) @ np.array([[c, s], [-s, c]]) + np.array([10.0, 3.0])
    a = L.pca_rotation_angle(p, t)
    assert np.isclose(np.mod(a - 0.7 + np.pi / 2, np.pi) - np.pi / 2, 0.0, atol=1e-9)
    assert L.pca_aligned(L.chamfer, p, t)[0] < 1e-12


def _frames(seed, s=3, n=30, grid=(40, 50)):
    rng = np.random.default_rng(seed)
    h, w = grid
    main = rng.uniform(0, 1, (n, 2)) * np.array([w, h])
    fr = np.mod(main[None] + rng.normal(0, 3.0, (s, n, 2)), np.array([w, h]))
    anchor = np.mod(main + rng.normal(0, 2.0, (n, 2)), np.array([w, h]))
    return fr, main, anchor, grid


@pytest.mark.parametrize("use_anchor", [False, True])
def test_frame_stability_matches_fd(use_anchor):
    fr, main, anchor, grid = _frames(7)
    an = anchor if use_anchor else None
    l, gf, gm, ga = L.frame_stability(grid, fr, main, an)
    np.testing.assert_allclose(gf, _fd(lambda x: L.frame_stability(grid, x, main, an)[0], fr),
                               atol=1e-7)
    np.testing.assert_allclose(gm, _fd(lambda x: L.frame_stability(grid, fr, x, an)[0], main),
                               atol=1e-7)
    if use_anchor:
        np.testing.assert_allclose(
            ga, _fd(lambda x: L.frame_stability(grid, fr, main, x)[0], anchor), atol=1e-7)
    else:
        assert ga is None


def test_frame_stability_closed_forms():
    """Every frame shifted by (a, b) from the main frame: each term is
    mean((a, b)^2) = (a^2 + b^2) / 2; a shift by whole periods is no
    displacement (minimum image); the frame-wise mean divides by the number
    of frames."""
    _, main, _, grid = _frames(8)
    h, w = grid
    a, b = 1.5, -2.0
    fr = np.stack([main + [a, b], main + [a, b]])
    assert L.frame_stability(grid, fr, main)[0] == pytest.approx((a * a + b * b) / 2)
    wrapped = np.stack([main + [w, -h], main + [2 * w, 0.0]])
    assert L.frame_stability(grid, wrapped, main)[0] == pytest.approx(0.0, abs=1e-24)
    two = np.stack([main + [a, 0.0], main])
    assert L.frame_stability(grid, two, main)[0] == pytest.approx(a * a / 2 / 2)
    # with the anchor equal to main the anchor term doubles the loss
    assert L.frame_stability(grid, fr, main, main)[0] == pytest.approx(a * a + b * b)


def test_smoothness_matches_fd_and_closed_form():
    a = np.random.default_rng(9).normal(size=12)
    l, g = L.smoothness(a)
    np.testing.assert_allclose(g, _fd(lambda x: L.smoothness(x)[0], a), atol=1e-8)
    assert L.smoothness(np.arange(5.0) * 0.3)[0] == pytest.approx(0.09)


@pytest.mark.parametrize("mode", ["dcd", "pointcloud"])
def test_folding_objective_matches_fd(mode):
    """The frame bookkeeping of the objective (evaluation, stability, main
    and anchor frames, weights) against finite differences in every frame's
    positions, centred branch (the PCA angle is stop_gradient, pinned on its
    own above)."""
    rng = np.random.default_rng(10)
    grid = (40, 50)
    base = rng.uniform(10, 30, (20, 2))
    pos = {t: base + rng.normal(0, 1.0, base.shape) for t in (3, 5, 6, 7, 8)}
    tgt = rng.normal(0, 4.0, (25, 2))
    kw = dict(loss_mode=mode, use_pca=False, stability_weight=0.7, max_dist_weight=0.4, alpha=0.5)
    ev, st, an = [3, 5], [6, 7, 8], 7

    def f(t, x):
        q = dict(pos)
        q[t] = x
        return L.folding_objective(q, tgt, grid, ev, st, an, **kw)[0]

    total, shape, stab, bars = L.folding_objective(pos, tgt, grid, ev, st, an, **kw)
    assert total > shape > 0.0 and stab > 0.0
    for t in pos:
        np.testing.assert_allclose(bars[t], _fd(lambda x: f(t, x), pos[t]), atol=2e-6,
                                   err_msg=f"frame {t}")
