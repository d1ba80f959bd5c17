"""The rod adjoint restatement (oracle/rod_adjoint.py) against central finite
differences of the float64 forward restatement.  CPU only."""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import rod_adjoint as A
from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import oracle_state


def _case(damping=0.5, cycles=4, seed=11, solver="pbd", steric=False):
    wl = W.small(grid=(32, 32), n_fil=2, n_nodes=12, seed=seed)
    rng = np.random.default_rng(seed)
    s = oracle_state(wl, np.float64)
    s = replace(s, bending_rest_angles=rng.normal(0.0, 0.2, s.bending_pairs.shape[0]),
                velocity=rng.normal(0.0, 0.05, s.position.shape))
    if solver == "cosserat":
        p = replace(R.COSSERAT_SCHEME_NO_FLUID, cycles_per_step=cycles, damping=damping,
                    enable_steric=steric, stiffness_stretch=5.0, config_dt=0.2)
    else:
        p = replace(R.PBD_SCHEME_NO_FLUID, cycles_per_step=cycles, damping=damping,
                    enable_steric=steric, stiffness_stretch=0.6, stiffness_shear=0.8)
    target = s.position + rng.normal(0.0, 1.0, s.position.shape)
    return s, p, target


def _fd(f, x, h):
    g = np.zeros_like(x)
    for i in range(x.size):
        e = np.zeros_like(x)
        e.flat[i] = h
        g.flat[i] = (f(x + e) - f(x - e)) / (2 * h)
    return g


@pytest.mark.parametrize("steric", [False, True])
@pytest.mark.parametrize("solver", ["pbd", "cosserat"])
def test_rollout_rest_angle_gradient_matches_fd(solver, steric):
    s, p, target = _case(solver=solver, steric=steric)
    steps = 6
    loss, g = A.rollout_grad(s, 1.0, p, steps, target)
    fd = _fd(lambda r: A.rollout_loss(s, 1.0, p, steps, target, rest=r), s.bending_rest_angles, 1e-6)
    assert np.isfinite(loss)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd))), (g, fd)
    cos = g @ fd / (np.linalg.norm(g) * np.linalg.norm(fd))
    assert cos > 0.999999


@pytest.mark.parametrize("steric", [False, True])
@pytest.mark.parametrize("solver", ["pbd", "cosserat"])
@pytest.mark.parametrize("which", ["position", "velocity", "theta"])
def test_step_vjp_directional_derivatives(which, solver, steric):
    """<vjp(w), u> equals the finite-difference derivative of <w, step(s)> in
    direction u, for the step inputs x, v and theta."""
    s, p, _ = _case(damping=0.7, cycles=3, seed=5, solver=solver, steric=steric)
    rng = np.random.default_rng(3)
    wx = rng.normal(size=s.position.shape)
    wv = rng.normal(size=s.position.shape)
    wt = rng.normal(size=s.theta.shape)
    xb, vb, tb, _ = A.step_vjp(s, 1.0, p, wx, wv, wt)
    u = rng.normal(size=getattr(s, which).shape)
    adj = {"position": xb, "velocity": vb, "theta": tb}[which]

    base = R.step(s, 1.0, p)
    h = 1e-6

    def f_unwrapped(eps):
        t = replace(s, **{which: getattr(s, which) + eps * u})
        o = R.step(t, 1.0, p)
        d = R.displacement(s.grid_shape, o.position, base.position)
        return (np.sum(wx * (base.position + d)) + np.sum(wv * o.velocity) + np.sum(wt * o.theta))

    fd = (f_unwrapped(h) - f_unwrapped(-h)) / (2 * h)
    assert abs(np.sum(adj * u) - fd) <= 1e-6 * max(1.0, abs(fd)), (np.sum(adj * u), fd)


def test_zero_damping_cuts_the_velocity_path():
    s, p, target = _case(damping=0.0)
    xb, vb, tb, rb = A.step_vjp(s, 1.0, p, np.ones_like(s.position), np.ones_like(s.position),
                                np.zeros_like(s.theta))
    assert np.all(vb == 0.0)


def test_rejects_fluid():
    s, p, target = _case()
    with pytest.raises(NotImplementedError):
        A.step_vjp(s, 1.0, replace(p, enable_fluid=True), s.position, s.position, s.theta)


def test_steric_vjp_directional_derivative():
    """The steric force's VJP (window term, bilinear-sample term and the
    deposit -> Gaussian filter -> sample path) against central differences of
    compute_field_steric_force_corrected."""
    s, p, _ = _case(seed=7, steric=True)
    rng = np.random.default_rng(9)
    g = rng.normal(size=s.position.shape)
    u = rng.normal(size=s.position.shape)
    out = A.steric_vjp(s.position, s.grid_shape, p, g)

    def f(eps):
        return np.sum(g * R.compute_field_steric_force_corrected(
            s.position + eps * u, s.grid_shape, p.steric_sigma, p.steric_neighbor_skip,
            p.steric_scatter_value)[2])

    h = 1e-6
    fd = (f(h) - f(-h)) / (2 * h)
    assert abs(np.sum(out * u) - fd) <= 1e-7 * max(1.0, abs(fd)), (np.sum(out * u), fd)


def test_folding_configuration_gradient_matches_fd():
    """The folding experiment's default physics (stable Cosserat, steric term,
    dt 0.02) on the 8 x 48-node folding workload, 30 steps: the rest-angle
    gradient along a random direction against a central difference with
    h = 1e-6 (CIC and bilinear kinks make larger steps and longer horizons
    cross pieces of the piecewise-smooth loss)."""
    wl, target = W.folding(steps=30)
    s = oracle_state(wl, np.float64)
    p = replace(R.COSSERAT_SCHEME_NO_FLUID, config_dt=W.FOLDING_DT, **W.FOLDING_COSSERAT)
    loss, g = A.rollout_grad(s, W.FOLDING_DT, p, 30, target)
    u = np.random.default_rng(9).normal(size=g.shape[0])
    h = 1e-6
    fp = A.rollout_loss(s, W.FOLDING_DT, p, 30, target, rest=s.bending_rest_angles + h * u)
    fm = A.rollout_loss(s, W.FOLDING_DT, p, 30, target, rest=s.bending_rest_angles - h * u)
    fd = (fp - fm) / (2 * h)
    assert abs(g @ u - fd) <= 1e-5 * abs(fd), (g @ u, fd)


@pytest.mark.parametrize("solver", ["pbd", "cosserat"])
def test_multi_frame_gradient_matches_fd(solver):
    """A loss over several frames (positions after 2, 4 and 6 steps, each
    against its own target) -- the shape of the folding experiment's
    objective, which scores evenly spaced frames: the adjoint with frame
    cotangents against central differences."""
    s, p, _ = _case(solver=solver, steric=True)
    rng = np.random.default_rng(21)
    targets = {t: s.position + rng.normal(0.0, 1.0, s.position.shape) for t in (2, 4, 6)}
    frames = {t: (lambda pos, tg=tg: A.msd_loss(pos, tg, s.grid_shape)) for t, tg in targets.items()}
    loss, g = A.rollout_frames_grad(s, 1.0, p, 6, frames)
    fd = _fd(lambda r: A.rollout_frames_loss(s, 1.0, p, 6, frames, rest=r), s.bending_rest_angles,
             1e-6)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd))), (g, fd)
    # the single-frame case is the final-state gradient
    l1, g1 = A.rollout_frames_grad(s, 1.0, p, 6, {6: frames[6]})
    l2, g2 = A.rollout_grad(s, 1.0, p, 6, targets[6])
    assert l1 == pytest.approx(l2) and np.allclose(g1, g2, rtol=1e-12, atol=1e-14)


def test_folding_raw_angle_gradient_matches_fd():
    """The folding loss of one initial state (shape over evaluation frames,
    stability against the main and anchor frames, smooth max, smoothness)
    through tanh(raw) * curvature_scale: its raw-angle gradient against
    central finite differences of the float64 rollout."""
    s, p, _ = _case()
    rng = np.random.default_rng(4)
    raw = rng.normal(0.0, 0.4, s.bending_pairs.shape[0])
    tgt = s.position[:16] + rng.normal(0.0, 1.5, (16, 2))
    frames = dict(eval_steps=[4, 6], stab_steps=[7, 8, 10], anchor_step=9)
    kw = dict(loss_mode="dcd", use_pca=False, stability_weight=0.8, max_dist_weight=0.5,
              alpha=0.3, smoothness_weight=0.2)
    total, g, _ = A.folding_raw_grad(s, 1.0, p, 10, raw, 0.6, tgt, **frames, **kw)
    fd = _fd(lambda r: A.folding_raw_grad(s, 1.0, p, 10, r, 0.6, tgt, **frames, **kw)[0], raw,
             1e-6)
    assert np.isfinite(total) and np.max(np.abs(fd)) > 1e-3
    np.testing.assert_allclose(g, fd, atol=1e-6 * max(1.0, np.max(np.abs(fd))))


@pytest.mark.parametrize("solver,steric", [("pbd", False), ("pbd", True), ("cosserat", False),
                                           ("cosserat", True)])
def test_rest_length_gradient_matches_fd(solver, steric):
    """Gradient of a two-frame loss with respect to the rest lengths (the
    folding experiment also optimises them, experiment.py:741-742) against
    central finite differences of the float64 rollout."""
    s, p, target = _case(solver=solver, steric=steric)
    s = replace(s, rest_lengths=s.rest_lengths * np.random.default_rng(3).uniform(
        0.8, 1.3, s.rest_lengths.shape))

    def objective(pos):
        l1, g1 = A.msd_loss(pos[4], target, s.grid_shape)
        l2, g2 = A.msd_loss(pos[8], target, s.grid_shape)
        return l1 + l2, {4: g1, 8: g2}

    def loss(lengths):
        q = replace(s, rest_lengths=lengths)
        return A.rollout_objective_grad(q, 1.0, p, 8, [4, 8], objective)[0]

    total, rg, lg, _ = A.rollout_objective_grad(s, 1.0, p, 8, [4, 8], objective, lengths=True)
    fd = _fd(loss, np.asarray(s.rest_lengths, np.float64), 1e-6)
    assert np.max(np.abs(fd)) > 1e-3
    np.testing.assert_allclose(lg, fd, atol=2e-6 * max(1.0, np.max(np.abs(fd))))


@pytest.mark.parametrize("solver", ["pbd", "cosserat"])
def test_folding_raw_length_gradient_matches_fd(solver):
    """The folding loss through the rest-length parameterisation lo +
    sigmoid(raw) (hi - lo) (experiment.py:741-742): raw-length gradient
    against central finite differences of the float64 rollout."""
    s, p, _ = _case(solver=solver)
    rng = np.random.default_rng(12)
    raw = rng.normal(0.0, 0.4, s.bending_pairs.shape[0])
    raw_l = rng.normal(-1.0, 0.3, s.pairs.shape[0])
    tgt = s.position[:16] + rng.normal(0.0, 1.5, (16, 2))
    kw = dict(eval_steps=[4, 6], stab_steps=[7, 8, 10], anchor_step=9, loss_mode="dcd",
              use_pca=False, stability_weight=0.8, max_dist_weight=0.5, alpha=0.3)
    total, g, gl, _ = A.folding_raw_grad(s, 1.0, p, 10, raw, 0.6, tgt, raw_lengths=raw_l, **kw)
    fd = _fd(lambda r: A.folding_raw_grad(s, 1.0, p, 10, raw, 0.6, tgt, raw_lengths=r, **kw)[0],
             raw_l, 1e-6)
    assert np.max(np.abs(fd)) > 1e-3
    np.testing.assert_allclose(gl, fd, atol=2e-6 * max(1.0, np.max(np.abs(fd))))
