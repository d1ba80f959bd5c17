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


# --------------------------------------------------------------------------
# Synthetic IB points (fluid.py:246-263, 364-368)
# --------------------------------------------------------------------------
def _chain_pairs(n):
    return np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int32)


def _syn_points(pos, pairs, grid, lattice, width, dtype=np.float32):
    pos = np.asarray(pos, dtype)
    sx = lattice[1] / grid[0]
    sy = lattice[0] / grid[1]
    gp = pos * np.array([sx, sy], dtype)
    return R.synthetic_ib_points(pos, pairs, grid, gp, width, 0.5 * (sx + sy))


@pytest.mark.parametrize("direction", ["+x", "+y", "-x"])
def test_synthetic_points_straight_chain_closed_form(direction):
    """A straight chain of spacing 2 (every tangent norm exact): the + point of
    each node lies at width * mean_scale along the tangent turned by +90
    degrees, the - point opposite; the halves are + points then - points."""
    grid, lattice, width = (32, 32), (64, 64), 0.75  # scale 2: offset 1.5 lattice cells
    n = 5
    t = {"+x": (1.0, 0.0), "+y": (0.0, 1.0), "-x": (-1.0, 0.0)}[direction]
    pos = np.array([[8.0 + 2.0 * k * t[0], 9.0 + 2.0 * k * t[1]] for k in range(n)])
    pts = _syn_points(pos, _chain_pairs(n), grid, lattice, width)
    normal = np.array([-t[1], t[0]])  # +90 degrees
    exp = np.concatenate([pos * 2.0 + 1.5 * normal, pos * 2.0 - 1.5 * normal])
    np.testing.assert_array_equal(pts, exp.astype(np.float32))


def test_synthetic_points_across_the_periodic_seam():
    """Edge vectors are minimum images: a chain crossing x = W has the same
    normal at every node as one that does not."""
    grid, lattice, width = (32, 32), (32, 32), 0.5
    pos = np.array([[28.0, 5.0], [30.0, 5.0], [0.0, 5.0], [2.0, 5.0]])
    pts = _syn_points(pos, _chain_pairs(4), grid, lattice, width)
    np.testing.assert_array_equal(pts[:4, 1], np.full(4, 5.5, np.float32))
    np.testing.assert_array_equal(pts[4:, 1], np.full(4, 4.5, np.float32))
    np.testing.assert_array_equal(pts[:4, 0], pos[:, 0].astype(np.float32))


def test_synthetic_points_invariants_random_graph():
    """Any graph: each offset has length width * mean_scale (up to the EPSIL in
    the norm), is normal to the summed incident edge vectors (computed here in
    float64 from the positions), lies on their left, and the - point mirrors
    the + point through the node; an isolated node keeps both on the node."""
    rng = np.random.default_rng(7)
    grid, lattice, width = (48, 40), (96, 80), 0.6
    n = 40
    pos = rng.uniform(0, 40, (n, 2)) * np.array([1.0, 48 / 40])
    pairs = np.array([[i, (i + 1 + rng.integers(0, 5)) % (n - 1)] for i in range(n - 1)], np.int32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    pts = _syn_points(pos, pairs, grid, lattice, width).astype(np.float64)
    scale = np.array([lattice[1] / grid[0], lattice[0] / grid[1]])
    ms = 0.5 * scale.sum()
    gp = pos.astype(np.float32).astype(np.float64) * scale
    size = np.array([grid[1], grid[0]], np.float64)
    tan = np.zeros((n, 2))
    for a, b in pairs:
        d = pos[b] - pos[a]
        d = (d + size / 2) % size - size / 2
        tan[a] += d
        tan[b] += d
    off = pts[:n] - gp
    # (float32 positions near 80: one ulp is 7.6e-6)
    np.testing.assert_allclose(pts[n:] - gp, -off, atol=3e-5)
    has = np.linalg.norm(tan, axis=1) > 1e-3
    assert has.sum() >= n - 2
    np.testing.assert_allclose(np.linalg.norm(off[has], axis=1), width * ms, atol=3e-5)
    ut = tan[has] / np.linalg.norm(tan[has], axis=1, keepdims=True)
    np.testing.assert_allclose((off[has] * ut).sum(1), 0.0, atol=3e-5)
    cross = ut[:, 0] * off[has, 1] - ut[:, 1] * off[has, 0]
    assert (cross > 0).all()
    iso = np.setdiff1d(np.arange(n), pairs.ravel())
    assert len(iso) > 0
    np.testing.assert_array_equal(pts[:n][iso], pts[n:][iso])


def test_synthetic_vanishing_width_equals_duplicated_nodes():
    """A width far below the positions' resolution puts both points on the
    node: the IB then equals the plain IB over the node set given twice, whose
    two copies feel identical forces, so their mean is either copy's force --
    bit for bit, lattice and u* included."""
    grid = lattice = (32, 32)
    p = _params(ibm_iterations=3)
    rng = np.random.default_rng(3)
    pos = np.array([[10.3 + 1.1 * k, 12.7 + 0.4 * k] for k in range(6)], np.float32)
    vel = rng.normal(0, 0.04, (6, 2)).astype(np.float32)
    f, _ = _rest_fluid(*lattice)
    f = f.astype(np.float32)
    v_s, u_s, f_s = R.immersed_boundary_interaction(
        pos, vel, grid, f, 1.0, replace(p, synthetic_node_width=1e-30), pairs=_chain_pairs(6))
    v_d, u_d, f_d = R.immersed_boundary_interaction(
        np.concatenate([pos, pos]), np.concatenate([vel, vel]), grid, f, 1.0, p)
    np.testing.assert_array_equal(v_s, v_d[:6])
    np.testing.assert_array_equal(v_d[:6], v_d[6:])
    np.testing.assert_array_equal(u_s, u_d)
    np.testing.assert_array_equal(f_s, f_d)


def test_synthetic_drag_is_the_mean_of_the_two_points():
    """Points far enough apart that the + and - stencils never share a cell:
    the synthetic IB equals the plain IB over the explicit 2N points (closed
    form of a straight chain), and node k's drag is the mean of the forces at
    points k and N + k."""
    grid, lattice = (32, 32), (32, 32)
    p = _params(ibm_iterations=4, ibm_relaxation=0.5)
    width = 3.0  # 6 cells between the + and - rows, K = 4
    n = 3
    pos = np.array([[10.0 + 2.0 * k, 14.25] for k in range(n)], np.float32)
    vel = np.array([[0.02, 0.05], [0.0, 0.04], [-0.01, 0.06]], np.float32)
    # a shear flow, so that the + and - rows see different fluid
    yy = np.arange(lattice[0])[:, None] + np.zeros(lattice)
    u0 = np.stack([0.03 * np.sin(2 * np.pi * yy / lattice[0]), np.zeros(lattice)])
    f = R.equilibrium(np.ones(lattice), u0).astype(np.float32)
    v_s, u_s, f_s = R.immersed_boundary_interaction(
        pos, vel, grid, f, 1.0, replace(p, synthetic_node_width=width), pairs=_chain_pairs(n))
    pts = np.concatenate([pos + [0.0, 3.0], pos - [0.0, 3.0]]).astype(np.float32)
    tr = R.IBTrace()
    _, u_e, f_e = R.immersed_boundary_interaction(pts, np.concatenate([vel, vel]), grid, f, 1.0,
                                                  p, trace=tr)
    np.testing.assert_array_equal(u_s, u_e)
    np.testing.assert_array_equal(f_s, f_e)
    F = tr.node_force.astype(np.float64)
    assert np.abs(F[:n] - F[n:]).max() > 1e-6  # the two sides differ: the mean matters
    v_exp = vel - ((F[:n] + F[n:]) / 2.0) / p.node_mass
    np.testing.assert_allclose(v_s, v_exp, rtol=0, atol=1e-7)


def test_synthetic_needs_pairs():
    grid = lattice = (16, 16)
    f, _ = _rest_fluid(*lattice)
    with pytest.raises(ValueError):
        R.immersed_boundary_interaction(np.zeros((2, 2)) + 4.0, np.zeros((2, 2)), grid, f, 1.0,
                                        _params(synthetic_node_width=0.5))
