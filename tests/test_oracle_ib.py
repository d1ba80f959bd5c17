"""Pins the oracle's immersed-boundary core (oracle/step_ref.py
``immersed_boundary_interaction``, after fluid.py:292-384) with closed forms.

Every case starts from fluid at rest, f = feq(1, 0), so rho is uniform and
u = 0.  A single node (or M nodes on one spot) then turns the multi-direct
forcing loop into a scalar recurrence that the mathematics fixes:

  u_corr(cell t) = a_t * A_n, U_n = S * A_n, A_{n+1} = A_n + clamp(r (V - S A_n))

with a_t the cell's share of each correction and S = sum_t w_t a_t:
  * one node (kw <= 1 everywhere, inv_kw = 1): a_t = w_t, S = sum w_t^2;
  * M coincident nodes: a_t = min(M w_t, 1) (the spread M c w_t times
    1 / max(M w_t, 1)), S = sum_t w_t min(M w_t, 1).
Without the clamps V - U_n = (1 - r S)^n V.  The drag each node feels is
F_n = sum_t w_t F_t with F_t = rho * postclamp(u_corr_t), and the node
velocity becomes v - F_n / (m * scale * dt).  u* = u + 0.5 F / rho.

Each test is chosen so that a plausible mistake fails it: a dropped or
inverted inv_kw, a relaxation applied twice or not at all, a skipped per-step
or post-accumulation clamp, a flipped drag sign, a dropped node mass, scale or
dt, the x scale taken from the wrong axis.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R

EPS = float(np.finfo(np.float32).eps)


def _rest_fluid(h, w, rho=1.0):
    f = R.equilibrium(np.full((h, w), rho), np.zeros((2, h, w)))
    return f, float(f.sum(0)[0, 0])


def _params(**kw):
    base = dict(ibm_kernel_size=4, ibm_iterations=5, ibm_relaxation=0.3, max_fluid_velocity=0.2,
                node_mass=1.0, viscosity=0.1, lambda_trt=0.1875, enable_force_clamp=False)
    base.update(kw)
    return replace(R.PBD_SCHEME, **base)


def _clamp(v, umax):
    mag = np.sqrt(float(v[0] * v[0] + v[1] * v[1]) + EPS)
    return v * min(1.0, umax / (mag + EPS))


def _run(pos, vel, grid, lattice, dt, p, rho0=1.0):
    f, rho = _rest_fluid(*lattice, rho=rho0)
    tr = R.IBTrace()
    v_new, ustar, f_new = R.immersed_boundary_interaction(
        np.asarray(pos, np.float64), np.asarray(vel, np.float64), grid, f, dt, p, trace=tr)
    return v_new, ustar, f_new, tr, rho


def _stencil(lattice_pos, lattice, K):
    flat, w = R.ib_stencil(np.asarray([lattice_pos], np.float64), lattice[0], lattice[1], K)
    return flat[0], w[0]


def test_single_node_geometric_series():
    """One node, no clamp engaged: V - U_n = (1 - r S)^n V with S = sum w^2."""
    grid = lattice = (32, 32)
    p = _params()
    V = np.array([0.05, -0.03])
    pos = np.array([10.3, 17.6])
    v_new, ustar, _, tr, rho = _run([pos], [V], grid, lattice, 1.0, p)
    flat, w = _stencil(pos, lattice, 4)
    S = float((w * w).sum())
    g = 1.0 - (1.0 - p.ibm_relaxation * S) ** p.ibm_iterations
    A = V * g / S                                   # sum of the five corrections
    # the force field: rho * w_t * A on the stencil, zero elsewhere
    F = np.zeros((2, lattice[0] * lattice[1]))
    F[:, flat] = rho * np.outer(A, w)
    np.testing.assert_allclose(tr.force.reshape(2, -1), F, rtol=0, atol=1e-15)
    # u* = u + 0.5 F / rho with u = 0
    np.testing.assert_allclose(ustar.reshape(2, -1), 0.5 * F / rho, rtol=0, atol=1e-15)
    # drag: F_n = rho V g, v_new = V - F_n / m
    np.testing.assert_allclose(v_new[0], V - rho * V * g, rtol=1e-12, atol=1e-15)
    # a relaxation applied twice or an extra spread factor changes g by far more
    g_wrong = 1.0 - (1.0 - 0.09 * S) ** 5
    assert abs(g - g_wrong) > 1e-3


@pytest.mark.parametrize("iters", [1, 3, 5, 8])
def test_single_node_deficit_decay_per_iteration(iters):
    """The interpolated deficit after n iterations, read back through the
    drag: F_n = rho (V - U_n) ... = rho V (1 - (1 - r S)^n)."""
    grid = lattice = (24, 24)
    p = _params(ibm_iterations=iters, ibm_relaxation=0.45, ibm_kernel_size=2)
    V = np.array([-0.02, 0.04])
    pos = np.array([7.55, 3.2])
    v_new, _, _, tr, rho = _run([pos], [V], grid, lattice, 1.0, p)
    _, w = _stencil(pos, lattice, 2)
    S = float((w * w).sum())
    g = 1.0 - (1.0 - 0.45 * S) ** iters
    np.testing.assert_allclose(tr.node_force[0], rho * V * g, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(v_new[0], V - rho * V * g, rtol=1e-12, atol=1e-16)


def test_mass_scale_dt_and_axis_quirk():
    """Lattice finer than the grid, non-square: scale_x = Wf / grid_shape[0]
    and scale_y = Hf / grid_shape[1] (fluid.py:225-226); the node target is
    v * scale * dt and the drag divides by node_mass * scale * dt, so
    v_new = V - rho V g / m; the stencil sits at position * scale."""
    grid = (16, 24)          # (H, W)
    lattice = (36, 48)       # (Hf, Wf): scale_x = 48 / 16 = 3, scale_y = 36 / 24 = 1.5
    dt, m = 0.5, 2.5
    p = _params(node_mass=m)
    V = np.array([0.02, -0.015])
    pos = np.array([5.2, 7.7])
    sx, sy = 48 / 16, 36 / 24
    v_new, ustar, _, tr, rho = _run([pos], [V], grid, lattice, dt, p)
    flat, w = _stencil(pos * [sx, sy], lattice, 4)
    # the support is the stencil of position * (3, 1.5): cells x in 14..17, y in 10..13
    nz = np.flatnonzero(np.abs(tr.force[0]).reshape(-1) > 0)
    assert set(nz.tolist()) == set(flat.tolist())
    cols = {int(c) % 48 for c in flat}
    rows = {int(c) // 48 for c in flat}
    assert cols == {14, 15, 16, 17} and rows == {10, 11, 12, 13}
    S = float((w * w).sum())
    g = 1.0 - (1.0 - p.ibm_relaxation * S) ** p.ibm_iterations
    vib = V * np.array([sx, sy]) * dt
    np.testing.assert_allclose(tr.node_force[0], rho * vib * g, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(v_new[0], V - rho * V * g / m, rtol=1e-12, atol=1e-16)


def _scalar_ib(V, a, w, n, r, umax):
    """The rank-one recurrence of the module docstring with both clamps;
    returns the per-cell force amplitude vector u_t (before rho)."""
    S = float((w * a).sum())
    A = np.zeros(2)
    for _ in range(n):
        c = _clamp(r * (V - S * A), umax)
        A = A + c
    u = np.array([_clamp(a_t * A, umax) for a_t in a])   # post-accumulation clamp per cell
    return A, u


def test_single_node_both_clamps_engaged():
    """A fast node: every per-step correction is clamped to |c| = umax and
    the accumulated u_corr exceeds umax on the central cells, so the post
    clamp engages there and not on the outer cells."""
    grid = lattice = (32, 32)
    p = _params(ibm_kernel_size=2)
    V = np.array([3.0, 1.0])
    pos = np.array([12.4, 20.1])
    v_new, ustar, _, tr, rho = _run([pos], [V], grid, lattice, 1.0, p)
    flat, w = _stencil(pos, lattice, 2)
    A, u = _scalar_ib(V, w, w, p.ibm_iterations, p.ibm_relaxation, 0.2)
    mags = np.linalg.norm(np.outer(w, A), axis=1)
    assert mags.max() > 0.2 and mags.min() < 0.2     # the post clamp engages on some cells only
    F = np.zeros((2, 32 * 32))
    F[:, flat] = rho * u.T
    np.testing.assert_allclose(tr.force.reshape(2, -1), F, rtol=0, atol=1e-14)
    Fn = (w[:, None] * (rho * u)).sum(0)
    np.testing.assert_allclose(v_new[0], V - Fn, rtol=1e-12, atol=1e-14)
    # without the per-step clamp the first correction alone would be 0.3 * |V| ~ 0.95
    assert np.linalg.norm(A) <= 5 * 0.2 + 1e-12
    # skipping the post clamp would leave the central cells above umax
    assert np.abs(np.linalg.norm(tr.u_corrected.reshape(2, -1)[:, flat], axis=0)).max() <= 0.2 + 1e-12


def test_force_clamp_option():
    """enable_force_clamp clamps |F| per cell to umax (fluid.py:332-335)."""
    grid = lattice = (32, 32)
    p = _params(enable_force_clamp=True, max_fluid_velocity=0.05)
    V = np.array([3.0, 1.0])
    pos = np.array([12.4, 20.1])
    _, _, _, tr, rho = _run([pos], [V], grid, lattice, 1.0, p)
    flat, w = _stencil(pos, lattice, 4)
    _, u = _scalar_ib(V, w, w, p.ibm_iterations, p.ibm_relaxation, 0.05)
    Fc = np.array([_clamp(rho * u_t, 0.05) for u_t in u])
    np.testing.assert_allclose(tr.force.reshape(2, -1)[:, flat], Fc.T, rtol=0, atol=1e-14)


@pytest.mark.parametrize("M", [1, 4, 9])
def test_coincident_nodes_kernel_weight_normalisation(M):
    """M nodes on one spot: kw_t = M w_t, each cell's correction is
    M c w_t / max(M w_t, 1), so a_t = min(M w_t, 1) and the deficit decays
    as (1 - r S_M)^n with S_M = sum w_t min(M w_t, 1).  Every node feels the
    whole drag F_n = rho V (1 - (1 - r S_M)^n)."""
    grid = lattice = (20, 20)
    p = _params(ibm_kernel_size=2)
    V = np.array([0.01, 0.02])
    pos = np.array([10.3, 6.8])
    v_new, ustar, _, tr, rho = _run([pos] * M, [V] * M, grid, lattice, 1.0, p)
    flat, w = _stencil(pos, lattice, 2)
    a = np.minimum(M * w, 1.0)
    if M == 4:
        assert (M * w > 1).any() and (M * w < 1).any()  # both branches of max(kw, 1)
    SM = float((w * a).sum())
    g = 1.0 - (1.0 - p.ibm_relaxation * SM) ** p.ibm_iterations
    A = V * g / SM
    np.testing.assert_allclose(tr.kernel_weight.reshape(-1)[flat], M * w, rtol=1e-14)
    F = np.zeros((2, 400))
    F[:, flat] = rho * np.outer(A, a)
    np.testing.assert_allclose(tr.force.reshape(2, -1), F, rtol=0, atol=1e-15)
    for k in range(M):
        np.testing.assert_allclose(v_new[k], V - rho * V * g, rtol=1e-12, atol=1e-16)
    # without inv_kw the decay factor would be (1 - r M S)^n instead
    if M > 1:
        S = float((w * w).sum())
        assert abs((1 - 0.3 * M * S) ** 5 - (1 - 0.3 * SM) ** 5) > 1e-3


def test_separated_nodes_do_not_interact():
    """Two nodes whose stencils share no cell each follow the single-node
    closed form with their own weights and velocities."""
    grid = lattice = (32, 32)
    p = _params()
    P = np.array([[5.3, 5.9], [20.6, 14.2]])
    V = np.array([[0.03, -0.01], [-0.02, 0.04]])
    v_new, _, _, tr, rho = _run(P, V, grid, lattice, 1.0, p)
    for k in range(2):
        _, w = _stencil(P[k], lattice, 4)
        S = float((w * w).sum())
        g = 1.0 - (1.0 - 0.3 * S) ** 5
        np.testing.assert_allclose(v_new[k], V[k] - rho * V[k] * g, rtol=1e-12, atol=1e-16)


def test_total_momentum_injected_matches_force_sum():
    """The lattice gains exactly sum_cells F of momentum in the collide-stream
    (Guo forcing, fluid.py:341-357), and the force field is the spread
    momentum of the single-node closed form: sum F = rho V g / S * sum w."""
    grid = lattice = (32, 32)
    p = _params()
    V = np.array([0.05, -0.03])
    pos = np.array([10.3, 17.6])
    f0, rho = _rest_fluid(*lattice)
    _, _, f_new, tr, _ = _run([pos], [V], grid, lattice, 1.0, p)
    e = R.LBM_VELOCITIES.astype(float)
    mom = np.einsum("ia,ixy->a", e, f_new) - np.einsum("ia,ixy->a", e, f0)
    _, w = _stencil(pos, lattice, 4)
    S = float((w * w).sum())
    g = 1.0 - (1.0 - 0.3 * S) ** 5
    np.testing.assert_allclose(mom, rho * V * g / S * w.sum(), rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(mom, tr.force.reshape(2, -1).sum(1), rtol=1e-12, atol=1e-16)
    # mass is conserved
    assert abs(f_new.sum() - f0.sum()) < 1e-11


def test_dense_fluid_force_carries_rho():
    """At rest density 1.3 the force is rho * (u_corr - u): the drag is
    1.3 V g, while u* = 0.5 F / rho = 0.5 u_corr does not see rho."""
    grid = lattice = (32, 32)
    p = _params()
    V = np.array([0.04, 0.01])
    pos = np.array([9.7, 11.2])
    v_new, ustar, _, tr, rho = _run([pos], [V], grid, lattice, 1.0, p, rho0=1.3)
    assert abs(rho - 1.3) < 1e-12
    flat, w = _stencil(pos, lattice, 4)
    S = float((w * w).sum())
    g = 1.0 - (1.0 - 0.3 * S) ** 5
    np.testing.assert_allclose(v_new[0], V - rho * V * g, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(ustar.reshape(2, -1)[:, flat], 0.5 * np.outer(V * g / S, w),
                               rtol=1e-12, atol=1e-16)
