"""Pins the CPU restatement (oracle/step_ref.py) before it is trusted as the
parity reference: the reference's own known-answer tests, closed forms,
invariants and brute force on tiny inputs, each chosen so that a plausible
mistake (a dropped term, a wrong sign or index, a transposed operand) fails."""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import oracle_params, oracle_state, run_oracle


# ---------------------------------------------------------------- utils KATs
def test_displacement_kat():
    """tests/test_utils.py:25-38."""
    d = R.displacement((10, 10), np.array([1.0, 1.0], np.float32), np.array([9.0, 9.0], np.float32))
    assert np.allclose(d, [2.0, 2.0])
    d2 = R.displacement((10, 10), np.array([5.0, 5.0], np.float32), np.array([4.0, 4.0], np.float32))
    assert np.allclose(d2, [1.0, 1.0])


def test_displacement_axis_order():
    """size (H, W) = (10, 20): x wraps by W = 20, y by H = 10."""
    d = R.displacement((10, 20), np.array([1.0, 1.0], np.float32), np.array([19.0, 9.0], np.float32))
    assert np.allclose(d, [2.0, 2.0])
    d = R.displacement((10, 20), np.array([1.0, 1.0], np.float32), np.array([9.0, 9.0], np.float32))
    assert np.allclose(d, [-8.0, 2.0])


def test_index_from_position_kat():
    """tests/test_utils.py:41-49."""
    iy, ix = R.index_from_position(np.array([[1.2, 3.8], [5.5, 0.1]], np.float32))
    assert list(iy) == [3, 0] and list(ix) == [1, 5]


def test_periodic_boundary_kat():
    """tests/test_utils.py:52-56."""
    w = R.periodic_boundary((10, 10), np.array([-1.0, 11.0], np.float32))
    assert np.allclose(w, [9.0, 1.0])


def test_gaussian_filter_kat():
    """tests/test_utils.py:59-68."""
    d = np.zeros((10, 10), np.float32)
    d[5, 5] = 1.0
    f = R.gaussian_filter(d, 1.0)
    assert int(np.argmax(f)) == 55
    assert np.allclose(f[5, 4], f[5, 6]) and np.allclose(f[4, 5], f[6, 5])


def _brute_gaussian(d, sigma):
    """Circular convolution with h = IDFT(exp(-2 pi^2 s^2 f^2)) per axis, by
    explicit sums in float64 (no FFT)."""
    H, W = d.shape

    def h1(n):
        k = np.arange(n)
        f = np.where(k < (n + 1) // 2, k, k - n) / n
        g = np.exp(-2 * np.pi ** 2 * sigma ** 2 * f ** 2)
        m = np.arange(n)
        return (g[None, :] * np.cos(2 * np.pi * np.outer(m, k) / n)).sum(1) / n

    hy, hx = h1(H), h1(W)
    out = np.zeros((H, W))
    for a in range(H):
        for b in range(W):
            if d[a, b] != 0:
                out += d[a, b] * np.outer(np.roll(hy, a), np.roll(hx, b))
    return out


@pytest.mark.parametrize("shape,sigma", [((12, 12), 2.0), ((9, 16), 1.5), ((10, 15), 1.0)])
def test_gaussian_filter_brute_force(shape, sigma):
    rng = np.random.default_rng(1)
    d = np.zeros(shape)
    idx = rng.integers(0, shape[0] * shape[1], 6)
    d.flat[idx] = rng.random(6)
    got = R.gaussian_filter(d, sigma)
    exp = _brute_gaussian(d, sigma)
    assert np.max(np.abs(got - exp)) < 1e-12
    assert abs(got.sum() - d.sum()) < 1e-12  # kernel(0) = 1: mass is conserved


@pytest.mark.parametrize("n", [4, 8, 16, 64, 128, 1024, 192, 96, 10, 50, 12, 21, 7, 11, 13, 30])
def test_stockham_fft_matches_pocketfft(n):
    """The restated device FFT network is a DFT: float64 agreement with numpy's
    pocketfft to rounding, forward and inverse, every radix path (4, 2, 3,
    generic 5 / 7 / 11 / 13)."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    re, im = R.stockham_fft(x.real.copy(), x.imag.copy())
    F = np.fft.fft(x, axis=-1)
    assert np.max(np.abs(re + 1j * im - F)) < 1e-12 * n
    re, im = R.stockham_fft(x.real.copy(), x.imag.copy(), inverse=True)
    assert np.max(np.abs(re + 1j * im - np.fft.ifft(x, axis=-1) * n)) < 1e-12 * n


@pytest.mark.parametrize("shape", [(48, 48), (64, 96), (10, 10), (192, 128)])
def test_gaussian_filter_restatement_matches_pocketfft(shape):
    rng = np.random.default_rng(7)
    d = rng.random(shape)
    a = R.gaussian_filter(d, 2.0)
    b = R.gaussian_filter_pocketfft(d, 2.0)
    assert np.max(np.abs(a - b)) < 1e-14
    a32 = R.gaussian_filter(d.astype(np.float32), 2.0)
    assert a32.dtype == np.float32
    assert np.linalg.norm(a32 - b) / np.linalg.norm(b) < 1e-6


@pytest.mark.parametrize("n,sigma", [(1024, 2.0), (256, 5.0), (48, 1.5), (10, 1.0), (33, 2.0)])
def test_gauss_taps_are_the_idft_kernel(n, sigma):
    """The direct filter's taps are the exact periodic kernel (brute-force
    inverse DFT of one transfer-function factor) on the kept offsets, the
    dropped tail holds at most CONV_TAIL of the mass, and the kernel is
    symmetric (the transfer function is even)."""
    k = np.arange(n)
    f = np.where(k < (n + 1) // 2, k, k - n) / n
    g = np.exp(-2 * np.pi ** 2 * sigma ** 2 * f ** 2)
    h = (g[None, :] * np.cos(2 * np.pi * np.outer(k, k) / n)).sum(1) / n
    lo, taps = R.gauss_taps(n, sigma)
    off = lo + np.arange(taps.size)
    assert np.max(np.abs(taps - h[off % n])) < 1e-13  # cos of large arguments in the brute force
    kept = np.zeros(n, bool)
    kept[off % n] = True
    assert np.abs(h[~kept]).sum() <= R.CONV_TAIL * np.abs(h).sum()
    assert taps.size == n or np.allclose(taps, taps[::-1], rtol=0, atol=1e-15)


def test_gauss_taps_long_kernel_falls_back():
    """sigma 1.5 on 512: the transfer function is cut at Nyquist at 1.5e-5,
    so the periodic kernel rings over the whole axis: FFT path."""
    assert R.gauss_taps(512, 1.5) is None
    d = np.random.default_rng(2).random((512, 512)).astype(np.float32)
    np.testing.assert_array_equal(R.gaussian_filter(d, 1.5), R.gaussian_filter_fft(d, 1.5))


@pytest.mark.parametrize("shape,sigma", [((48, 48), 2.0), ((64, 96), 1.5), ((10, 10), 1.0),
                                         ((256, 128), 5.0)])
def test_direct_filter_matches_pocketfft(shape, sigma):
    """The direct filter against numpy's FFT filter: in float64 arithmetic to
    the dropped tail (2 CONV_TAIL of the data's scale), in float32 to float32
    rounding."""
    rng = np.random.default_rng(5)
    d = rng.random(shape)
    b = R.gaussian_filter_pocketfft(d, sigma)
    a64 = R.gaussian_filter_direct(d, sigma)
    assert np.max(np.abs(a64 - b)) <= 2 * R.CONV_TAIL * np.abs(d).max() + 1e-14
    a32 = R.gaussian_filter_direct(d.astype(np.float32), sigma)
    assert a32.dtype == np.float32
    assert np.linalg.norm(a32 - b) / np.linalg.norm(b) < 1e-6


def test_fftfreq_matches_numpy():
    for n in (8, 9, 10, 128, 192):
        assert np.allclose(R.fftfreq32(n, np.float64), np.fft.fftfreq(n))


# ---------------------------------------------------------------- steric
def test_update_steric_potential_kat():
    """tests/test_forces.py:7-32: node at (5,5) on 10x10, defaults sigma 12.5, value 8."""
    _, st = R.update_steric_potential(np.array([[5.0, 5.0]], np.float32), (10, 10), 12.5, 8.0)
    assert st[5, 5] > 0 and st[5, 6] > 0


def test_cic_deposit_brute_force():
    rng = np.random.default_rng(2)
    H, W = 9, 13
    pos = (rng.random((40, 2)) * [W, H] * 1.3 - 2).astype(np.float32)
    dens, _ = R.update_steric_potential(pos, (H, W), 2.0, 1.5)
    exp = np.zeros((H, W), np.float32)
    pw = np.mod(pos, np.array([W, H], np.float32))
    corners = []
    for n in range(pos.shape[0]):
        x, y = pw[n]
        fx, fy = np.floor(x), np.floor(y)
        x0, y0 = int(fx) % W, int(fy) % H
        dx, dy = x - fx, y - fy
        corners.append((x0, (x0 + 1) % W, y0, (y0 + 1) % H, dx, dy))
    one = np.float32(1.0)
    sv = np.float32(1.5)
    for q in range(4):  # corner-major, node-minor (forces.py:112-115)
        for (x0, x1, y0, y1, dx, dy) in corners:
            w = [(one - dx) * (one - dy) * sv, dx * (one - dy) * sv,
                 (one - dx) * dy * sv, dx * dy * sv][q]
            cy, cx = [(y0, x0), (y0, x1), (y1, x0), (y1, x1)][q]
            exp[cy, cx] = np.float32(exp[cy, cx] + np.float32(w))
    assert np.array_equal(dens, exp)
    assert abs(float(dens.astype(np.float64).sum()) - 40 * 1.5) < 1e-4


def test_force_interp_matches_scipy():
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(4)
    V = rng.random((11, 14))
    pos = rng.random((50, 2)) * [14, 11]
    for dy, dx in ((0, 0.5), (0, -0.5), (0.5, 0), (-0.5, 0)):
        got = R.sample_bilinear_wrap(V, pos[:, 1] + dy, pos[:, 0] + dx)
        exp = map_coordinates(V, [pos[:, 1] + dy, pos[:, 0] + dx], order=1, mode="grid-wrap")
        assert np.max(np.abs(got - exp)) < 1e-12


def test_force_interp_linear_field_exact():
    a, b = 0.3, -0.7
    yy, xx = np.mgrid[0:20, 0:20].astype(np.float64)
    V = a * xx + b * yy
    pos = np.array([[5.3, 7.9], [10.01, 3.5], [14.6, 14.2]])
    f = R.force_interp(pos, V)
    assert np.allclose(f, [[-a, -b]] * 3, atol=1e-12)


def test_steric_window_two_nodes_closed_form():
    sigma, sv = 2.0, 1.0
    pos = np.array([[10.0, 10.0], [11.5, 9.0]])
    corr = R.steric_window_subtraction(pos, (32, 32), sigma, 5, sv)
    r = pos[0] - pos[1]
    K = np.exp(-(r @ r) / (2 * sigma ** 2)) / (2 * np.pi * sigma ** 2)
    assert np.allclose(corr[0], sv / sigma ** 2 * K * r, rtol=1e-14)
    assert np.allclose(corr[1], -corr[0], rtol=1e-14)
    # window 0: no subtraction
    assert np.all(R.steric_window_subtraction(pos, (32, 32), sigma, 0, sv) == 0)


def test_steric_force_self_consistent_isolated_node():
    """An isolated node's own bump is removed by the window: residual force is
    only the FD/bilinear discretisation error, far below the raw force scale."""
    pos = np.array([[16.3, 15.7]])
    _, st, f = R.compute_field_steric_force_corrected(pos, (32, 32), 2.0, 5, 1.0)
    raw = R.force_interp(pos, st)
    assert np.linalg.norm(f) < 0.2 * np.linalg.norm(raw) + 1e-3


# ---------------------------------------------------------------- rods
def _line_state(n, spacing=1.0, theta=0.0, grid=(64, 64), rest=0.0):
    pairs, bp, _ = __import__("microcosmos_amd.topology", fromlist=["x"]).line_instances([n])
    pos = np.stack([20 + np.arange(n) * spacing * np.cos(theta),
                    30 + np.arange(n) * spacing * np.sin(theta)], 1)
    return R.State(position=pos, velocity=np.zeros_like(pos), debug_vector=pos.copy(),
                   pairs=pairs, theta=np.full(n - 1, theta), rest_lengths=np.full(n - 1, spacing),
                   bending_pairs=bp, bending_rest_angles=np.full(bp.shape[0], rest),
                   bending_stiffness=np.full(bp.shape[0], 0.5), grid_shape=grid,
                   steric=np.zeros(grid), fluid_velocity=np.zeros((2,) + grid),
                   f_grid=np.zeros((9,) + grid))


@pytest.mark.parametrize("fn,p", [(R.pbd_rod_constraint, R.PBD_SCHEME),
                                  (R.stable_cosserat_constraint, R.COSSERAT_SCHEME)])
def test_straight_rest_chain_is_fixed_point(fn, p):
    s = _line_state(12, theta=0.4)
    p0, t0 = s.position.copy(), s.theta.copy()
    for _ in range(5):
        fn(s, p)
    assert np.max(np.abs(s.position - p0)) < 1e-12
    assert np.max(np.abs(s.theta - t0)) < 1e-12


def test_pbd_single_edge_closed_form():
    s = _line_state(2, spacing=1.0)
    s.position = np.array([[10.0, 10.0], [11.6, 10.0]])  # stretched to L = 1.6
    p = replace(R.PBD_SCHEME, stiffness_stretch=0.6)
    R.pbd_rod_constraint(s, p)
    L = np.linalg.norm(s.position[1] - s.position[0])
    assert abs(L - (1.6 - 0.6 * 0.6)) < 1e-12           # each end moves k_st/2 (L - l0)
    assert np.allclose(s.position.mean(0), [10.8, 10.0])  # momentum-free correction


def test_pbd_one_hinge_theta_update():
    s = _line_state(3)
    s.theta = np.array([0.1, 0.5])
    s.bending_rest_angles = np.array([0.2])
    s.bending_stiffness = np.array([0.5])
    t_in, t_out = 0.1, 0.5
    corr = ((t_out - t_in) - 0.2) * 0.5 * 0.5
    R.pbd_rod_constraint(s, R.PBD_SCHEME)
    assert np.allclose(s.theta, [t_in + 1.8 * corr, t_out - 1.8 * corr], atol=1e-15)


def test_theta_wrap_crossing_pi():
    s = _line_state(3)
    s.theta = np.array([3.0, -3.0])  # true difference 2 pi - 6 = 0.283
    s.bending_rest_angles = np.array([0.0])
    s.bending_stiffness = np.array([1.0])
    R.pbd_rod_constraint(s, R.PBD_SCHEME)
    d = 2 * np.pi - 6.0
    assert np.allclose(s.theta, [3.0 + 1.8 * d * 0.5, -3.0 - 1.8 * d * 0.5], atol=1e-14)


def test_cosserat_single_edge_closed_form():
    s = _line_state(2, spacing=2.0)
    s.position = np.array([[10.0, 10.0], [12.5, 10.0]])
    s.debug_vector = s.position.copy()
    p = replace(R.COSSERAT_SCHEME, stiffness_stretch=50.0, config_dt=0.05)
    R.stable_cosserat_constraint(s, p)
    w, k, l = 400.0, 50.0, 2.0
    c = 2.5 / l - 1.0                       # r/l - d along x
    dx = (k / l) * c / (w + k / l ** 2)     # src moves +, tgt moves -
    assert np.allclose(s.position, [[10.0 + dx, 10.0], [12.5 - dx, 10.0]], atol=1e-12)


def _cos_line(n, rest_length=1.0):
    s = _line_state(n, spacing=rest_length, grid=(64, 64))
    s.position = np.stack([np.arange(n) * rest_length, np.full(n, 32.0)], 1).astype(np.float32)
    s = s.astype(np.float32)
    s.bending_stiffness = np.ones(s.bending_pairs.shape[0], np.float32)
    return s


def _cos_steps(s, n):
    p = R.COSSERAT_SCHEME_NO_FLUID
    for _ in range(n):
        s = R.step(s, 0.05, p)
    return s


def test_reference_cosserat_invariants():
    """tests/test_stable_cosserat.py:81-154 run through the oracle."""
    s0 = _cos_line(10)
    s1 = _cos_steps(s0, 20)
    assert np.abs(s1.position - s0.position).max() < 0.01
    s = _cos_line(10)
    s.position = s.position + 0.3
    s = _cos_steps(s, 30)
    r = R.displacement((64, 64), s.position[s.pairs[:, 1]], s.position[s.pairs[:, 0]])
    assert np.abs(np.linalg.norm(r, axis=-1) - s.rest_lengths).max() < 0.15
    s = _cos_line(5, 2.0)
    s.bending_rest_angles = np.full(3, np.pi / 4, np.float32)
    s = _cos_steps(s, 50)
    d = np.mod(s.theta[s.bending_pairs[:, 1]] - s.theta[s.bending_pairs[:, 0]] + np.pi, 2 * np.pi) - np.pi
    assert np.abs(d - np.pi / 4).mean() < 0.3
    s = _cos_line(3, 3.0)
    s.bending_rest_angles = np.array([np.pi / 2], np.float32)
    s = _cos_steps(s, 100)
    d = np.mod(s.theta[1] - s.theta[0] + np.pi, 2 * np.pi) - np.pi
    assert abs(d) > 0.3


# ---------------------------------------------------------------- IB stencil
def _brute_stencil(x, y, h, w, K):
    half = (K - 1) // 2
    hw = K / 2.0
    sx, sy = np.floor(x) - half, np.floor(y) - half
    taps = []
    for dx in range(K):          # outer
        for dy in range(K):      # inner
            nx, ny = sx + dx, sy + dy
            phi = lambda r: 0.25 * (1 + np.cos(np.pi * abs(r) / hw / 2)) if abs(r) / hw <= 2 else 0.0
            taps.append((int(ny) % h * w + int(nx) % w, phi(nx - x) * phi(ny - y)))
    tot = sum(t[1] for t in taps)
    return [t[0] for t in taps], [t[1] / (tot + float(R.EPSIL)) for t in taps]


@pytest.mark.parametrize("K", [2, 3, 4, 6])
def test_stencil_brute_force(K):
    rng = np.random.default_rng(K)
    pos = rng.random((30, 2)) * 40 - 5
    flat, wts = R.ib_stencil(pos, 24, 20, K)
    for n in range(pos.shape[0]):
        idx, w = _brute_stencil(pos[n, 0], pos[n, 1], 24, 20, K)
        assert list(flat[n]) == idx
        assert np.allclose(wts[n], w, rtol=1e-14, atol=1e-16)


def test_stencil_k4_weights_at_integer_position():
    """K = 4 on frac = 0: taps at -1..2 with cos(pi |d| / 4) weights (the
    reference's kernel is twice the textbook Peskin width)."""
    _, w = R.ib_stencil(np.array([[10.0, 10.0]]), 32, 32, 4)
    w1 = np.array([0.25 * (1 + np.cos(np.pi * d / 4)) for d in (1, 0, 1, 2)])
    w1 = w1 / w1.sum()
    assert np.allclose(w.reshape(4, 4), np.outer(w1, w1), atol=1e-7)
    assert np.allclose(w1, [0.2661, 0.3118, 0.2661, 0.1559], atol=1e-4)


def test_stencil_partition_and_adjointness():
    rng = np.random.default_rng(9)
    h, w = 20, 24
    pos = rng.random((25, 2)) * [w, h]
    flat, wts = R.ib_stencil(pos, h, w, 4)
    assert np.allclose(wts.sum(1), 1.0, atol=1e-6)
    const = np.ones((2, h, w)) * 0.7
    assert np.allclose(R.ib_interpolate(flat, wts, const), 0.7 * wts.sum(1)[:, None], atol=1e-14)
    U = rng.normal(size=(2, h, w))
    C = rng.normal(size=(25, 2))
    lhs = (R.ib_interpolate(flat, wts, U) * C).sum()
    rhs = (U * R.ib_spread(flat, wts, C, h, w)).sum()
    assert abs(lhs - rhs) < 1e-12


# ---------------------------------------------------------------- LBM
def test_feq_moments():
    rng = np.random.default_rng(0)
    rho = 1 + 0.1 * rng.random((6, 7))
    u = 0.05 * rng.normal(size=(2, 6, 7))
    f = R.equilibrium(rho, u)
    assert np.allclose(f.sum(0), rho, atol=1e-14)
    mom = np.einsum("ia,ixy->axy", R.LBM_VELOCITIES.astype(float), f)
    assert np.allclose(mom, rho * u, atol=1e-14)


def test_lbm_uniform_fixed_point_and_streaming_permutation():
    rho = np.ones((8, 9))
    u = np.zeros((2, 8, 9))
    f = R.equilibrium(rho, u)
    op, om = R.trt_omegas(0.1, 0.1875, np.dtype(np.float64))
    fp, _ = R.collide_trt_guo(f, rho, u, np.zeros((2, 8, 9)), op, om)
    assert np.allclose(fp, f, atol=1e-15)
    rng = np.random.default_rng(2)
    g = rng.random((9, 8, 9)) + 0.1
    s = R.stream(g)
    for i, (ex, ey) in enumerate(R.LBM_VELOCITIES):
        assert np.array_equal(s[i], np.roll(g[i], (ey, ex), axis=(0, 1)))
    # a population at (y, x) moves to (y + ey, x + ex)
    g = np.full((9, 8, 9), 0.5)
    g[5, 3, 4] = 2.0
    assert R.stream(g)[5, 4, 5] == 2.0


def test_trt_guo_mass_and_momentum():
    rng = np.random.default_rng(3)
    rho = 1 + 0.05 * rng.random((10, 12))
    u = 0.02 * rng.normal(size=(2, 10, 12))
    f = R.equilibrium(rho, u) + 1e-3 * rng.normal(size=(9, 10, 12))
    rho, u = R.macroscopic(f)
    F = 1e-3 * rng.normal(size=(2, 10, 12))
    op, om = R.trt_omegas(1 / 30, 0.1875, np.dtype(np.float64))
    fp, _ = R.collide_trt_guo(f, rho, u, F, op, om)
    assert np.allclose(fp.sum(0), f.sum(0), atol=1e-14)
    dm = (np.einsum("ia,ixy->axy", R.LBM_VELOCITIES.astype(float), fp)
          - np.einsum("ia,ixy->axy", R.LBM_VELOCITIES.astype(float), f))
    assert np.allclose(dm, F, atol=1e-14)


def test_trt_omegas():
    op, om = R.trt_omegas(0.1, 0.1875, np.dtype(np.float64))
    assert abs(op - 1.25) < 1e-15
    assert abs((1 / op - 0.5) * (1 / om - 0.5) - 0.1875) < 1e-14
    op, om = R.trt_omegas(0.001, 0.25, np.dtype(np.float64))
    assert op == 1.9  # clipped (fluid.py:343)


def test_taylor_green_shear_wave_decay():
    """Viscous decay of u_x = U sin(k y): amplitude ~ exp(-nu k^2 t)."""
    H, W = 64, 8
    nu = 0.1
    U = 1e-3
    k = 2 * np.pi / H
    y = np.arange(H)[:, None] * np.ones((1, W))
    u = np.zeros((2, H, W))
    u[0] = U * np.sin(k * y)
    f = R.equilibrium(np.ones((H, W)), u)
    op, om = R.trt_omegas(nu, 0.1875, np.dtype(np.float64))
    T = 200
    for _ in range(T):
        rho, uu = R.macroscopic(f)
        fp, _ = R.collide_trt_guo(f, rho, uu, np.zeros_like(uu), op, om)
        f = R.stream(fp)
    _, uu = R.macroscopic(f)
    amp = 2 * np.mean(uu[0] * np.sin(k * y))
    assert abs(amp / U - np.exp(-nu * k * k * T)) < 2e-3


# ---------------------------------------------------------------- coupled step
def test_reference_worm_behaviour():
    """tests/test_worm.py:15-152 through the oracle: a sine swimmer moves."""
    n, grid, dt = 20, (50, 50), 0.01
    pairs, bp, _ = __import__("microcosmos_amd.topology", fromlist=["x"]).line_instances([n])
    x = np.linspace(5.0, 45.0, n)
    amp = 15.0
    kk = 2 * np.pi * 2 / (x[-1] - x[0])
    pos = np.stack([x, 25 + amp * np.sin(kk * x)], 1).astype(np.float32)
    d = R.displacement(grid, pos[pairs[:, 1]], pos[pairs[:, 0]])
    th = np.arctan2(d[:, 1], d[:, 0]).astype(np.float32)

    def rest(phase):
        t = np.arctan(amp * kk * np.cos(kk * x + phase))
        return (t[1:-1] - t[:-2]).astype(np.float32)

    w = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36], np.float32)
    s = R.State(position=pos, velocity=np.zeros_like(pos), debug_vector=np.zeros_like(pos),
                pairs=pairs, theta=th, rest_lengths=np.full(n - 1, 2.0, np.float32),
                bending_pairs=bp, bending_rest_angles=rest(0.0),
                bending_stiffness=np.full(n - 2, 0.5, np.float32), grid_shape=grid,
                steric=np.zeros(grid, np.float32), fluid_velocity=np.zeros((2,) + grid, np.float32),
                f_grid=np.broadcast_to(w[:, None, None], (9,) + grid).copy())
    com0 = s.position[:, 0].mean()
    phase = 0.0
    for _ in range(200):
        s = replace(s, bending_rest_angles=rest(phase))
        s = R.step(s, dt, R.PBD_SCHEME)
        phase += 0.1
    assert np.all(np.isfinite(s.position)) and np.all(np.isfinite(s.velocity))
    disp = abs(com0 - s.position[:, 0].mean())
    assert 0.2 < disp < 50


def test_plain_case_fp32_tracks_fp64():
    """The 128^2 plain case stays finite and the FP32 restatement tracks its
    FP64 run (the error reference) over 20 steps."""
    wl = W.plain_128()
    p = oracle_params(wl)
    a = run_oracle(wl, 20, p)
    b = run_oracle(wl, 20, p, dtype=np.float64)
    assert np.all(np.isfinite(a.f_grid))
    assert np.linalg.norm(a.f_grid - b.f_grid) / np.linalg.norm(b.f_grid) < 1e-6
    du = np.linalg.norm(a.fluid_velocity - b.fluid_velocity) / np.linalg.norm(b.fluid_velocity)
    assert du < 1e-3


@pytest.mark.parametrize("n,sigma", [(1024, 2.0), (256, 2.0), (48, 2.0), (32, 1.5), (10, 1.0),
                                     (96, 1.5), (262, 2.0), (256, 5.0), (512, 1.5), (4096, 2.0)])
def test_library_filter_taps_equal_the_restatement(n, sigma):
    """The plan's host tabulation of the direct filter (mc_filter_taps, no
    device work) gives the restatement's offsets and float32 taps bit for bit,
    and the same refusal where the FFT is needed."""
    import ctypes
    from microcosmos_amd import _lib
    lib = _lib.load()
    lo = ctypes.c_int32()
    buf = (ctypes.c_float * 256)()
    cnt = lib.mc_filter_taps(n, sigma, ctypes.byref(lo), buf, 256)
    ref = R.gauss_taps(n, sigma)
    if ref is None:
        assert cnt == 0
        return
    assert cnt == ref[1].size and lo.value == ref[0]
    np.testing.assert_array_equal(np.frombuffer(buf, np.float32)[:cnt], ref[1].astype(np.float32))


def _f32_math_inputs():
    rng = np.random.default_rng(11)
    special = np.array([0.0, -0.0, np.nan, np.inf, -np.inf, 8192.0, -8192.0, 8192.001, 1e-30,
                        -1e-30, 1e-40, np.pi / 4, -np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2,
                        2 * np.pi, 1e4, -1e6, 3.0e38], dtype=np.float32)
    return np.concatenate([special, rng.uniform(-8200, 8200, 200_000).astype(np.float32),
                           rng.uniform(-12, 12, 200_000).astype(np.float32),
                           np.linspace(-4, 4, 40_001, dtype=np.float32)])


def test_sincos_f32_accuracy():
    """The float32 sine / cosine of the filament cycles and IB weights against
    float64 numpy: absolute error below 1e-7 for |x| <= 8192 (reduction by a
    three-part pi/2 split, Taylor series to x^9 / x^10), within 2 ulp where
    the value is not near a zero of the function, sin(-x) = -sin(x),
    cos(-x) = cos(x), and the float64 fallback beyond 8192."""
    x = _f32_math_inputs()
    x = x[np.isfinite(x)]
    s, c = R.sincos_f32(x)
    xd = x.astype(np.float64)
    ts, tc = np.sin(xd), np.cos(xd)
    assert np.abs(s - ts).max() < 1e-7 and np.abs(c - tc).max() < 1e-7
    for got, ref in ((s, ts), (c, tc)):
        far = (np.abs(ref) > 0.05) & (np.abs(xd) < 16)
        ulp = np.spacing(np.abs(ref.astype(np.float32))).astype(np.float64)
        assert (np.abs(got[far] - ref[far]) / ulp[far]).max() <= 2.0
    s2, c2 = R.sincos_f32(-x)
    np.testing.assert_array_equal(s2, -s)
    np.testing.assert_array_equal(c2, c)
    big = np.abs(x) > 8192
    np.testing.assert_array_equal(s[big], ts[big].astype(np.float32))
    # quadrants: sin / cos at the multiples of pi/2 to 1e-7
    q = np.arange(-8, 9, dtype=np.float32) * np.float32(np.pi / 2)
    sq, cq = R.sincos_f32(q)
    np.testing.assert_allclose(sq, np.sin(q.astype(np.float64)), atol=1e-7)
    np.testing.assert_allclose(cq, np.cos(q.astype(np.float64)), atol=1e-7)


def test_exp_f32_accuracy():
    """The float32 exp of the steric window against float64 numpy: within
    1.5 ulp for normal results, one rounding of the denormal range, +0 below
    -104, NaN propagates, exp(0) = 1 exactly."""
    rng = np.random.default_rng(12)
    x = np.concatenate([rng.uniform(-110, 0, 400_000), rng.uniform(-8, 8, 200_000),
                        rng.uniform(0, 88, 50_000), -np.linspace(0, 30, 30_001)]).astype(np.float32)
    e = R.exp_f32(x)
    t = np.exp(x.astype(np.float64))
    normal = t >= np.finfo(np.float32).tiny
    ulp = np.spacing(t[normal].astype(np.float32)).astype(np.float64)
    assert (np.abs(e[normal] - t[normal]) / ulp).max() <= 1.5
    assert (np.abs(e[~normal] - t[~normal]) / 2.0 ** -149).max() <= 1.0
    assert R.exp_f32(np.float32(0.0)) == 1.0
    assert np.isnan(R.exp_f32(np.float32(np.nan)))
    assert R.exp_f32(np.float32(-1e6)) == 0.0 and R.exp_f32(np.float32(-105.0)) == 0.0


def test_library_f32_math_equals_the_restatement():
    """The kernels' float32 sine / cosine / exp (mc_math_f32: the same source
    evaluated on the host, no device work) equal the restatement bit for bit,
    special values included."""
    import ctypes
    from microcosmos_amd import _lib
    lib = _lib.load()
    x = _f32_math_inputs()
    s = np.empty_like(x)
    c = np.empty_like(x)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib.mc_math_f32(0, ptr(x), x.size, ptr(s), ptr(c)) == 0
    rs, rc = R.sincos_f32(x)
    np.testing.assert_array_equal(s.view(np.uint32), rs.view(np.uint32))
    np.testing.assert_array_equal(c.view(np.uint32), rc.view(np.uint32))
    xe = np.concatenate([x, -np.abs(x), np.linspace(-110, 0, 50_001, dtype=np.float32)])
    e = np.empty_like(xe)
    assert lib.mc_math_f32(1, ptr(xe), xe.size, ptr(e), None) == 0
    np.testing.assert_array_equal(e.view(np.uint32), R.exp_f32(xe).view(np.uint32))
