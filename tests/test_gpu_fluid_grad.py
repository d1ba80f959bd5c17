"""Device adjoint of the fluid-coupled step (mc_rod_adjoint_forward /
_backward_params with enable_fluid): the recorded forward equals the plain
rollout bit for bit, and directional derivatives of the rest-angle and
rest-length gradients agree with central differences of the float64 oracle
(the reference's jax.grad through simulate with the IB positions
stop_gradient'ed, fluid.py:239)."""
from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import device_state, oracle_state

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


def _case(seed=5, steric=False, K=4, grid=(40, 40)):
    wl = W.small(grid=grid, n_fil=2, n_nodes=12, seed=seed)
    rng = np.random.default_rng(seed)
    wl.bending_rest_angles = rng.normal(0.0, 0.3, wl.bending_pairs.shape[0]).astype(np.float32)
    wl.velocity = rng.normal(0.0, 0.05, wl.position.shape).astype(np.float32)
    over = dict(wl.config_overrides)
    over.update(cycles_per_step=4, ibm_kernel_size=K, enable_steric=steric,
                max_fluid_velocity=0.2, viscosity=0.05)
    return wl, over


def _device(wl, over, steps, dev, with_lengths=True):
    from microcosmos_amd import PBD_SCHEME
    from microcosmos_amd.grad import RodAdjoint
    nodes, edges, fields = device_state(wl, dev)
    ad = RodAdjoint(nodes, edges, fields, replace(PBD_SCHEME, **over), wl.dt)
    x = ad.forward(steps, 1)
    return ad, x


def _oracle_final(wl, over, steps, rest=None, lengths=None, ib=None):
    """The float64 oracle after ``steps`` steps; ``ib`` a list: empty, it
    records each step's IB stencil positions, filled, it holds them fixed
    (the reference's stop_gradient, fluid.py:239)."""
    s = oracle_state(wl, np.float64)
    if rest is not None:
        s = replace(s, bending_rest_angles=rest)
    if lengths is not None:
        s = replace(s, rest_lengths=lengths)
    p = replace(R.PBD_SCHEME, **over)
    record = ib is not None and len(ib) == 0
    for k in range(steps):
        if record:
            s = R.step(s, wl.dt, p, ib_trace=ib)
        else:
            s = R.step(s, wl.dt, p, ib_position=None if ib is None else ib[k])
    return s


def test_fluid_forward_equals_rollout(dev):
    from microcosmos_amd import PBD_SCHEME, Rollout
    wl, over = _case()
    steps = 7
    _say(f"fluid adjoint forward vs rollout, {steps} steps")
    ad, x = _device(wl, over, steps, dev)
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, replace(PBD_SCHEME, **over), wl.dt)
    ro.run(steps, use_graph=False)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(x), _np(ro.position))
    np.testing.assert_array_equal(_np(ad.velocity), _np(ro.velocity))
    np.testing.assert_array_equal(_np(ad.f_grid), _np(ro.f_grid))


@pytest.mark.parametrize("steric,K", [(False, 4), (False, 2), (True, 4)])
def test_fluid_gradient_vs_finite_differences(dev, steric, K):
    wl, over = _case(steric=steric, K=K)
    steps = 6
    _say(f"fluid adjoint vs central FD, steric={steric} K={K}, {steps} steps")
    ad, x = _device(wl, over, steps, dev)
    rng = np.random.default_rng(3)
    c = rng.normal(size=wl.position.shape)
    pb = torch.as_tensor(c.astype(np.float32), device=dev)
    rg, lg, *_ = ad.backward_params({steps: pb})
    torch.cuda.synchronize()
    rg, lg = _np(rg).astype(np.float64), _np(lg).astype(np.float64)
    ib = []
    base = _oracle_final(wl, over, steps, ib=ib)
    assert np.abs(R.displacement(wl.grid_shape, _np(x).astype(np.float64), base.position)).max() < 1e-3

    def loss(rest=None, lengths=None):
        s = _oracle_final(wl, over, steps, rest, lengths, ib=ib)
        return float(np.sum(c * R.displacement(wl.grid_shape, s.position, base.position)))

    a0 = wl.bending_rest_angles.astype(np.float64)
    l0 = wl.rest_lengths.astype(np.float64)
    h = 1e-5
    u = rng.normal(size=a0.shape)
    fd_a = (loss(rest=a0 + h * u) - loss(rest=a0 - h * u)) / (2 * h)
    w = rng.normal(size=l0.shape)
    fd_l = (loss(lengths=l0 + h * w) - loss(lengths=l0 - h * w)) / (2 * h)
    _say(f"  rest angles: adjoint {rg @ u:.6e}  FD {fd_a:.6e}")
    _say(f"  rest lengths: adjoint {lg @ w:.6e}  FD {fd_l:.6e}")
    assert abs(rg @ u - fd_a) <= 2e-3 * abs(fd_a) + 1e-6
    assert abs(lg @ w - fd_l) <= 2e-3 * abs(fd_l) + 1e-6
    # the fluid's share of the gradient is well above that tolerance: the
    # rod-only gradient of the same rollout differs
    from microcosmos_amd import PBD_SCHEME_NO_FLUID
    from microcosmos_amd.grad import RodAdjoint
    nodes, edges, fields = device_state(wl, dev)
    ad0 = RodAdjoint(nodes, edges, fields, replace(PBD_SCHEME_NO_FLUID, **{**over, "enable_fluid": False}), wl.dt)
    ad0.forward(steps, 1)
    rg0 = _np(ad0.backward_params({steps: pb})[0]).astype(np.float64)
    _say(f"  rod-only adjoint {rg0 @ u:.6e}")
    assert abs(rg0 @ u - fd_a) > 2e-2 * abs(fd_a)


def test_reference_differentiability_scenario(dev):
    """tests/test_differentiability.py:51-71: PBD_SCHEME (fluid on, K = 2),
    a 10-node loop of rest length 3 on a 100 x 100 grid at rest, dt 0.01; the
    gradient of mean(position^2) with respect to the rest angles and rest
    lengths after 1 step (as the reference) and 2 steps (the fluid acting):
    finite, and equal to central differences of the float64 oracle.  After 4
    steps only finiteness is asserted, as the reference does: the loop sits
    at its equilibrium, where velocities ((x' - x) / dt of positions near 50,
    speeds near 1e-5) carry float32 rounding of the same size, so the float32
    trajectory is no longer the float64 one the differences linearise."""
    from microcosmos_amd import PBD_SCHEME
    from microcosmos_amd.grad import RodAdjoint
    from microcosmos_amd.structs import Edges, Fields, Nodes
    from microcosmos_amd.topology import compute_bending_pairs, make_edges
    n, shape, dt = 10, (100, 100), 0.01
    pairs, dep = make_edges(n, "loop")
    radius = 3.0 / (2.0 * np.sin(np.pi / n))
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False)
    pos = (np.array(shape, np.float64) / 2.0 + radius * np.stack([np.cos(ang), np.sin(ang)], 1))
    pos = pos.astype(np.float32)
    d = R.displacement(shape, pos[pairs[:, 1]], pos[pairs[:, 0]])
    theta = np.arctan2(d[:, 1], d[:, 0]).astype(np.float32)
    bp, _ = compute_bending_pairs(pairs, dep, n)
    rest = np.full(bp.shape[0], 2.0 * np.pi / n, np.float32)
    lens = np.full(pairs.shape[0], 3.0, np.float32)
    wts = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36], np.float32)
    f0 = np.broadcast_to(wts[:, None, None], (9,) + shape).astype(np.float32)
    for steps in (1, 2, 4):
        _say(f"reference differentiability scenario, {steps} step(s)")
        t = lambda a, dt_=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt_, device=dev)
        nodes = Nodes(position=t(pos), velocity=t(np.zeros_like(pos)), color_bend=t(np.arange(n) % 3, torch.int32),
                      debug_vector=t(np.zeros_like(pos)))
        edges = Edges(pairs=t(pairs, torch.int32), theta=t(theta), rest_lengths=t(lens),
                      bending_pairs=t(bp, torch.int32), bending_rest_angles=t(rest),
                      bending_stiffness=t(np.full(bp.shape[0], 0.5, np.float32)))
        fields = Fields(grid_shape=shape, steric=t(np.zeros(shape, np.float32)),
                        fluid_velocity=t(np.zeros((2,) + shape, np.float32)), f_grid=t(f0))
        ad = RodAdjoint(nodes, edges, fields, PBD_SCHEME, dt)
        x = ad.forward(steps, 1)
        xb = 2.0 * x / (2 * n)  # d mean(x^2) / dx
        rg, lg, *_ = ad.backward_params({steps: xb})
        torch.cuda.synchronize()
        rg, lg = _np(rg).astype(np.float64), _np(lg).astype(np.float64)
        assert rg.shape == rest.shape and lg.shape == lens.shape
        assert np.isfinite(rg).all() and np.isfinite(lg).all()
        if steps == 4:
            continue
        s0 = R.State(position=pos.astype(np.float64), velocity=np.zeros((n, 2)),
                     debug_vector=np.zeros((n, 2)), pairs=pairs, theta=theta.astype(np.float64),
                     rest_lengths=lens.astype(np.float64), bending_pairs=bp,
                     bending_rest_angles=rest.astype(np.float64),
                     bending_stiffness=np.full(bp.shape[0], 0.5), grid_shape=shape,
                     steric=np.zeros(shape), fluid_velocity=np.zeros((2,) + shape),
                     f_grid=f0.astype(np.float64))

        ib = []
        sb = s0
        for _ in range(steps):
            sb = R.step(sb, dt, R.PBD_SCHEME, ib_trace=ib)

        def loss(ra=None, ln=None):
            s = replace(s0, bending_rest_angles=rest.astype(np.float64) if ra is None else ra,
                        rest_lengths=lens.astype(np.float64) if ln is None else ln)
            for k in range(steps):
                s = R.step(s, dt, R.PBD_SCHEME, ib_position=ib[k])
            return float(np.mean(s.position ** 2))

        rng = np.random.default_rng(steps)
        h = 1e-6
        u, w = rng.normal(size=rest.shape), rng.normal(size=lens.shape)
        fd_a = (loss(ra=rest + h * u) - loss(ra=rest - h * u)) / (2 * h)
        fd_l = (loss(ln=lens + h * w) - loss(ln=lens - h * w)) / (2 * h)
        _say(f"  angles {rg @ u:.6e} vs {fd_a:.6e}; lengths {lg @ w:.6e} vs {fd_l:.6e}")
        assert abs(rg @ u - fd_a) <= 2e-3 * abs(fd_a) + 1e-5
        assert abs(lg @ w - fd_l) <= 2e-3 * abs(fd_l) + 1e-5
