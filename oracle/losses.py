"""Folding losses of the reference's filament-folding experiment, with their
gradients with respect to the particle positions (CPU restatement).

TEST INFRASTRUCTURE ONLY, like ``step_ref``: imported by ``tests/`` as the
checker of the device losses, never by the product path.

experiments/filament_folding/experiment.py:
  _chamfer_loss      :57-63   mean_i min_j d2 + mean_j min_i d2
  _dcd_loss          :105-128 the same terms divided by the density count of
                              the chosen nearest point (counts carry no gradient)
  _smooth_max_nn     :139-149 logsumexp(alpha sqrt(min_j d2 + 1e-8)) / alpha
  _pca_rotation_angle:66-84   principal-axis angle difference (no gradient)
  _rotate_particles  :87-92   rotation of the centred particle cloud
  *_pca_aligned               min over the angle and the angle + pi, targets centred
  _frame_stab        :878-893 mean min-image squared displacement of a frame
                              from the main (and anchor) frame
  loss_fn            :803-916 shape loss over the evaluation frames + weighted
                              stability and smooth-max terms

Gradients follow jax.grad's rules: the derivative of a min over ties is split
equally among the tied entries (jax's reduce-min / minimum rule), argmin takes
the first index, and stop_gradient quantities are constants.  Every function
is float64.  Pinned by central finite differences and by closed forms in
``tests/test_losses_oracle.py``; the frame indices of the folding objective
are pinned there by hand-evaluated values of the reference's formulas at its
configured sizes.
"""
from __future__ import annotations

import numpy as np


def _d2(p, t):
    diff = p[:, None, :] - t[None, :, :]
    return (diff * diff).sum(axis=-1)


def _min_grad_weights(d2, axis):
    """Weights of jax's min-reduction derivative: 1/k on each of the k tied
    minima along ``axis``."""
    m = d2.min(axis=axis, keepdims=True)
    ind = (d2 == m).astype(np.float64)
    return ind / ind.sum(axis=axis, keepdims=True)


def chamfer(p, t):
    """experiment.py:57-63; returns (loss, d loss / d p)."""
    p = np.asarray(p, np.float64)
    t = np.asarray(t, np.float64)
    n, m = p.shape[0], t.shape[0]
    d2 = _d2(p, t)
    loss = d2.min(axis=1).mean() + d2.min(axis=0).mean()
    w = _min_grad_weights(d2, 1) / n + _min_grad_weights(d2, 0) / m  # (N, M)
    g = 2.0 * (w.sum(axis=1)[:, None] * p - w @ t)
    return float(loss), g


def dcd(p, t):
    """experiment.py:105-128; returns (loss, d loss / d p)."""
    p = np.asarray(p, np.float64)
    t = np.asarray(t, np.float64)
    n, m = p.shape[0], t.shape[0]
    d2 = _d2(p, t)
    nn_p = d2.argmin(axis=1)
    cnt_p = np.zeros(m)
    np.add.at(cnt_p, nn_p, 1.0)
    nn_t = d2.argmin(axis=0)
    cnt_t = np.zeros(n)
    np.add.at(cnt_t, nn_t, 1.0)
    cp = cnt_p[nn_p]            # (N,)
    ct = cnt_t[nn_t]            # (M,)
    loss = (d2.min(axis=1) / cp).mean() + (d2.min(axis=0) / ct).mean()
    w = _min_grad_weights(d2, 1) / (n * cp[:, None]) + _min_grad_weights(d2, 0) / (m * ct[None, :])
    g = 2.0 * (w.sum(axis=1)[:, None] * p - w @ t)
    return float(loss), g


def smooth_max_nn(p, t, alpha=0.1):
    """experiment.py:139-149; returns (loss, d loss / d p)."""
    p = np.asarray(p, np.float64)
    t = np.asarray(t, np.float64)
    d2 = _d2(p, t)
    dm = d2.min(axis=1)
    r = np.sqrt(dm + 1e-8)
    z = alpha * r
    zmax = z.max()
    e = np.exp(z - zmax)
    loss = (zmax + np.log(e.sum())) / alpha
    s = e / e.sum()                            # d loss / d r_i = s_i
    dr = s / (2.0 * r)                         # d loss / d dm_i
    w = _min_grad_weights(d2, 1) * dr[:, None]
    g = 2.0 * (w.sum(axis=1)[:, None] * p - w @ t)
    return float(loss), g


def pca_rotation_angle(p, t):
    """experiment.py:66-84: angle of the targets' principal axis minus that of
    the particles' (eigenvector of the larger eigenvalue of each 2x2
    covariance + 1e-6 I)."""
    def axis(x):
        c = x - x.mean(axis=0)
        cov = c.T @ c / c.shape[0] + 1e-6 * np.eye(2)
        _, v = np.linalg.eigh(cov)
        return v[:, -1]
    vp, vt = axis(np.asarray(p, np.float64)), axis(np.asarray(t, np.float64))
    return float(np.arctan2(vt[1], vt[0]) - np.arctan2(vp[1], vp[0]))


def rotate(p, angle):
    """experiment.py:87-92: rotation of the centred particles."""
    c = p - p.mean(axis=0)
    co, si = np.cos(angle), np.sin(angle)
    return np.stack([c[:, 0] * co - c[:, 1] * si, c[:, 0] * si + c[:, 1] * co], axis=-1)


def pca_aligned(loss_fn, p, t, angle=None, **kw):
    """The *_pca_aligned wrappers (experiment.py:95-102, 131-136, 152-160):
    min of the loss at the angle and at the angle + pi, with the targets
    centred; the angle carries no gradient.  The gradient goes through the
    rotation and the centring (minimum ties split equally, as jnp.minimum)."""
    p = np.asarray(p, np.float64)
    t = np.asarray(t, np.float64)
    a = pca_rotation_angle(p, t) if angle is None else angle
    tc = t - t.mean(axis=0)
    out = []
    for ang in (a, a + np.pi):
        q = rotate(p, ang)
        l, gq = loss_fn(q, tc, **kw)
        co, si = np.cos(ang), np.sin(ang)
        gc = np.stack([gq[:, 0] * co + gq[:, 1] * si, -gq[:, 0] * si + gq[:, 1] * co], axis=-1)
        out.append((l, gc - gc.mean(axis=0)))
    (l0, g0), (l1, g1) = out
    if l0 < l1:
        return l0, g0
    if l1 < l0:
        return l1, g1
    return l0, 0.5 * (g0 + g1)


def min_image(grid_shape, a, b):
    """utils.py:18-23: floor-mod minimum-image displacement a - b, with the
    size flipped to (W, H) for (x, y) positions."""
    h, w = grid_shape
    size = np.array([float(w), float(h)])
    return np.mod(np.asarray(a, np.float64) - b + size / 2.0, size) - size / 2.0


def frame_stability(grid_shape, frames, main, anchor=None):
    """experiment.py:878-893: mean over the frames of mean(d(x_s, main)^2)
    (+ mean(d(x_s, anchor)^2) when the anchor is used).  The floor-mod's
    derivative is 1 in its first argument (jax's rem rule).  Returns (loss,
    d/d frames (S, n, 2), d/d main, d/d anchor or None)."""
    frames = np.asarray(frames, np.float64)
    s = frames.shape[0]
    k = frames[0].size
    loss = 0.0
    g = np.zeros_like(frames)
    gm = np.zeros_like(frames[0])
    ga = np.zeros_like(frames[0]) if anchor is not None else None
    for i in range(s):
        dm = min_image(grid_shape, frames[i], main)
        loss += (dm * dm).mean()
        g[i] += 2.0 * dm / (k * s)
        gm -= 2.0 * dm / (k * s)
        if anchor is not None:
            da = min_image(grid_shape, frames[i], anchor)
            loss += (da * da).mean()
            g[i] += 2.0 * da / (k * s)
            ga -= 2.0 * da / (k * s)
    return loss / s, g, gm, ga


def smoothness(angles):
    """experiment.py:895-896: mean((a[1:] - a[:-1])^2) over the rest angles;
    returns (loss, d loss / d angles)."""
    a = np.asarray(angles, np.float64)
    d = a[1:] - a[:-1]
    g = np.zeros_like(a)
    g[1:] += 2.0 * d / d.size
    g[:-1] -= 2.0 * d / d.size
    return float((d * d).mean()), g


def _centred(loss_fn, p, t, **kw):
    """The non-PCA branch (experiment.py:826-827): both clouds centred; the
    gradient goes through the particles' centring."""
    p = np.asarray(p, np.float64)
    t = np.asarray(t, np.float64)
    l, g = loss_fn(p - p.mean(axis=0), t - t.mean(axis=0), **kw)
    return l, g - g.mean(axis=0)


def folding_objective(positions, target, grid_shape, eval_steps, stab_steps, anchor_step,
                      loss_mode="dcd", use_pca=True, stability_weight=1.0, max_dist_weight=1.0,
                      alpha=0.1):
    """experiment.py:803-916 for the point-cloud modes, as a function of the
    frame positions: ``positions`` maps a step count to the positions after
    that many steps (the reference's timeseries index + 1); ``eval_steps``
    and ``stab_steps`` are the scored frames, the main frame is the last
    evaluation frame and ``anchor_step`` the anchor (None: not used).  The
    smoothness term depends on the rest angles only and is not part of this
    function.  Returns (total, shape, stability, {step: d total / d
    positions})."""
    fn = chamfer if loss_mode == "pointcloud" else dcd
    bars = {t: np.zeros_like(np.asarray(positions[t], np.float64)) for t in positions}
    ne = len(eval_steps)

    def frame(loss, p, **kw):
        return pca_aligned(loss, p, target, **kw) if use_pca else _centred(loss, p, target, **kw)

    shape = 0.0
    for t in eval_steps:
        l, g = frame(fn, positions[t])
        shape += l / ne
        bars[t] += g / ne
    main = eval_steps[-1]
    if stab_steps:
        fr = np.stack([np.asarray(positions[t], np.float64) for t in stab_steps])
        stab, g, gm, ga = frame_stability(grid_shape, fr, positions[main],
                                          None if anchor_step is None else positions[anchor_step])
        for i, t in enumerate(stab_steps):
            bars[t] += stability_weight * g[i]
        bars[main] += stability_weight * gm
        if anchor_step is not None:
            bars[anchor_step] += stability_weight * ga
    else:
        stab = 0.0   # the last frame is the main frame: zero displacement
    md = 0.0
    if max_dist_weight > 0.0:
        for t in eval_steps:
            l, g = frame(smooth_max_nn, positions[t], alpha=alpha)
            md += l / ne
            bars[t] += max_dist_weight * g / ne
    return shape + stability_weight * stab + max_dist_weight * md, shape, stab, bars
