"""CPU restatement (NumPy) of the reference's coupled fluid-filament time step.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and nowhere else.  The product path (``microcosmos_amd``) never
imports, calls or falls back to it.

It restates, operation for operation and in the same floating-point evaluation
order, the JAX code of ``ProjectProgramAMark/microcosmos`` for the path

    simulate.step -> steric FFT field -> PBD / Cosserat rod projection
                  -> immersed boundary (IB) + D2Q9 TRT lattice Boltzmann

Every function cites the reference file:line it follows.  The module is
dtype-generic: run it on float32 arrays for the parity oracle and on float64
arrays for the error reference (the reference itself never enables x64).

Evaluation-order conventions (chosen to be reproducible on the GPU):
  * every ``.at[].add`` scatter is an in-order fold from zero (``np.add.at``
    applies updates in index order);
  * every small reduction (9 populations, 16 stencil taps, 2n+1 steric window)
    is a left fold in index order -- numpy's pairwise ``sum`` is never used;
  * in float32 runs, the filament angles' sin/cos, the IB weights' cos and the
    steric window's exp use float32 routines (``sincos_f32``, ``exp_f32``:
    Cody-Waite reduction and Taylor polynomials, restated operation by
    operation, within ~2 ulp / 1e-7 absolute of the float64 functions);
    atan2 and the other exps are evaluated in float64 and rounded to float32.
    The device does the same.  XLA-CPU's own float32 polynomials are not
    reproduced (they differ from either by a few ulp);
  * the 2-D FFT is restated as the device's Stockham radix-4/2/3 network
    (rows, then columns, per-axis inverse scaling); it is pinned against
    numpy's pocketfft in float64 and a brute-force DFT.  jnp.fft (ducc, FP32)
    rounds differently; that difference is not reproduced.

Parity status per function (how each is pinned; see tests/test_oracle_*.py):
  displacement, periodic_boundary, index_from_position  -- reference KATs
      (tests/test_utils.py:25-56)
  gaussian_filter        -- reference peak/symmetry KAT (tests/test_utils.py:59-68),
                            float64 brute-force circular convolution, mass
  update_steric_potential -- reference positivity KAT (tests/test_forces.py:7-32),
                            brute-force CIC, mass = N * scatter_value
  force_interp           -- scipy.ndimage.map_coordinates(order=1, mode="grid-wrap"),
                            exact gradient of a linear field
  steric window          -- closed form for two isolated nodes
  pbd_rod_constraint     -- straight rest chain fixed point, single-edge closed form,
                            hand-derived one-hinge theta update
  stable_cosserat        -- straight rest chain fixed point, single-edge closed form,
                            reference invariants (tests/test_stable_cosserat.py:81-154)
  stencil (IB)           -- brute-force integer taps, partition of unity,
                            constant interpolation, spread/interp adjointness
  IB forcing loop, post-clamp, force, drag (immersed_boundary_interaction)
                         -- closed forms at rest fluid (tests/test_oracle_ib.py):
                            one node's deficit (1 - r S)^n, M coincident nodes'
                            (1 - r S_M)^n through inv_kw, both clamps through the
                            rank-one recurrence, mass / scale / dt / axis quirk,
                            rho in the force, injected momentum
  synthetic_ib_points and the 2N-point IB with the averaged drag
                         -- closed forms of straight chains (+x, +y, -x) and of
                            a chain across the periodic seam; offset length,
                            normality and side on a random branched graph; a
                            vanishing width equal to the plain IB over the nodes
                            given twice; separated points equal to the plain IB
                            over the explicit 2N points with the mean drag
  LBM                    -- feq moments (float64), mass conservation, uniform fixed
                            point, exact streaming permutation, Guo momentum input,
                            Taylor-Green viscous decay exp(-2 nu k^2 t)
  step                   -- reference worm behaviour (tests/test_worm.py:15-152)
  resize_linear          -- identity at equal sizes, partition of unity, exact
                            linear interpolation of a ramp when upsampling,
                            closed-form 2x downsampling taps (1, 3, 3, 1) / 8
  group_com / swim tracker -- float64 mean, translation by a lattice period,
                            a blob straddling the boundary, brute-force
                            per-step cumulative sum
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

# utils.py:9 -- float32 machine epsilon, used as a numpy float32 scalar so that
# float32 arrays stay float32 and float64 arrays stay float64.
EPSIL = np.float32(np.finfo(np.float32).eps)

# fluid.py:15-23 -- D2Q9 lattice, direction order and weights.
LBM_VELOCITIES = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                           [1, 1], [-1, 1], [-1, -1], [1, -1]])
LBM_WEIGHTS = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])
OPP_INDICES = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])


# --------------------------------------------------------------------------
# Parameters and state
# --------------------------------------------------------------------------
@dataclass(frozen=True)
class Params:
    """Numeric fields of ConstraintSolverConfig (solver/config.py:9-71).

    ``config_dt`` is ConstraintSolverConfig.dt (the Cosserat inertia step);
    the step's own ``dt`` is passed to ``step`` separately, as in the reference.
    """
    solver_type: str = "pbd"
    # ConstraintSolverConfig.bending_constraint as a name, "pbd" or
    # "cosserat" (None: the one solver_type names)
    bending_constraint: Optional[str] = None
    cycles_per_step: int = 20
    stiffness_stretch: float = 0.6
    stiffness_shear: float = 1.0
    damping: float = 1.0
    enable_fluid: bool = True
    viscosity: float = 0.01
    max_fluid_velocity: float = 0.2
    config_dt: float = 0.05
    ibm_iterations: int = 5
    synthetic_node_width: float = 0.0
    ibm_kernel_size: int = 3
    ibm_relaxation: float = 0.3
    node_mass: float = 1.0
    lambda_trt: float = 0.25
    enable_force_clamp: bool = False
    enable_steric: bool = False
    steric_strength: float = 1.0
    steric_sigma: float = 1.5
    steric_neighbor_skip: int = 5
    steric_scatter_value: float = 1.0


# solver/config.py:74-97
PBD_SCHEME = Params(solver_type="pbd", cycles_per_step=20, stiffness_stretch=0.6,
                    stiffness_shear=1.0, damping=1.0, enable_fluid=True, viscosity=0.01,
                    max_fluid_velocity=0.2, ibm_iterations=5, ibm_kernel_size=2)
# solver/config.py:100-119
PBD_SCHEME_NO_FLUID = Params(solver_type="pbd", cycles_per_step=10, stiffness_stretch=0.7,
                             stiffness_shear=0.7, damping=0.0, enable_fluid=False,
                             viscosity=0.1, max_fluid_velocity=0.1)
# solver/config.py:126-144
COSSERAT_SCHEME_NO_FLUID = Params(solver_type="cosserat", cycles_per_step=20,
                                  stiffness_stretch=50.0, stiffness_shear=50.0, damping=0.0,
                                  config_dt=0.05, enable_fluid=False, viscosity=0.1,
                                  max_fluid_velocity=0.1, steric_strength=100.0)
# solver/config.py:152-171
COSSERAT_SCHEME = Params(solver_type="cosserat", cycles_per_step=20, stiffness_stretch=50.0,
                         stiffness_shear=50.0, damping=1.0, config_dt=0.05, enable_fluid=True,
                         viscosity=0.01, max_fluid_velocity=0.2, ibm_iterations=5,
                         ibm_kernel_size=2)


@dataclass
class State:
    """Nodes + Edges + Fields of the reference (structs/nodes.py:5-15,
    structs/edges.py:6-21, structs/fields.py:5-16), as numpy arrays."""
    position: np.ndarray          # (N, 2) x, y
    velocity: np.ndarray          # (N, 2)
    debug_vector: np.ndarray      # (N, 2)
    pairs: np.ndarray             # (E, 2) int32 (src, tgt)
    theta: np.ndarray             # (E,)
    rest_lengths: np.ndarray      # (E,)
    bending_pairs: np.ndarray     # (B, 2) int32 (edge_in, edge_out)
    bending_rest_angles: np.ndarray  # (B,)
    bending_stiffness: np.ndarray    # (B,)
    grid_shape: tuple             # (H, W) of the node domain
    steric: np.ndarray            # (H, W)
    fluid_velocity: np.ndarray    # (2, Hf, Wf)
    f_grid: np.ndarray            # (9, Hf, Wf)
    color_bend: Optional[np.ndarray] = None

    def copy(self) -> "State":
        return replace(self, **{k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                                for k, v in self.__dict__.items()})

    def astype(self, dtype) -> "State":
        fl = ("position", "velocity", "debug_vector", "theta", "rest_lengths",
              "bending_rest_angles", "bending_stiffness", "steric", "fluid_velocity", "f_grid")
        return replace(self, **{k: np.asarray(getattr(self, k), dtype=dtype) for k in fl})


# --------------------------------------------------------------------------
# Elementwise helpers
# --------------------------------------------------------------------------
def _tf(fn, x):
    """Transcendental in the array's precision; float32 via a float64 evaluation."""
    x = np.asarray(x)
    if x.dtype == np.float32:
        return fn(x.astype(np.float64)).astype(np.float32)
    return fn(x)


# Float32 sine / cosine / exp of the device's latency-bound chains (the
# filament cycles, the IB weights, the steric window), restated operation by
# operation: Cody-Waite reduction by pi/2 (three-part split, exact products for
# |x| <= 8192) and the Taylor series of sin to x^9, cos to x^10 on
# [-pi/4, pi/4]; exp by ln 2 reduction (two-part split), the series to x^7 and
# ldexp.  Every step is one IEEE float32 operation in the device's order, so
# the device reproduces these bit for bit.  Pinned against the float64 library
# functions (absolute error, ulp bounds) in the oracle tests.
_H = float.fromhex
_F = np.float32
_TWO_OVER_PI = _F(_H("0x1.45f306p-1"))
_PIO2 = (_F(_H("0x1.92p+0")), _F(_H("0x1.fb4p-12")), _F(_H("0x1.4442d2p-24")))
_SIN_C = (_F(_H("-0x1.555556p-3")), _F(_H("0x1.111112p-7")), _F(_H("-0x1.a01a02p-13")),
          _F(_H("0x1.71de3ap-19")))
_COS_C = (_F(_H("0x1.555556p-5")), _F(_H("-0x1.6c16c2p-10")), _F(_H("0x1.a01a02p-16")),
          _F(_H("-0x1.27e4fcp-22")))
_LOG2E = _F(_H("0x1.715476p+0"))
_LN2 = (_F(_H("0x1.63p-1")), _F(_H("-0x1.bd0106p-13")))
_EXP_C = (_F(1.0), _F(0.5), _F(_H("0x1.555556p-3")), _F(_H("0x1.555556p-5")),
          _F(_H("0x1.111112p-7")), _F(_H("0x1.6c16c2p-10")), _F(_H("0x1.a01a02p-13")))
_SINCOS_LIMIT = _F(8192.0)


def sincos_f32(x):
    """(sin x, cos x) in float32 by the device's reduction and polynomials;
    beyond |x| = 8192 the float64 functions rounded to float32."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        k = np.rint(x * _TWO_OVER_PI)
        r = ((x - k * _PIO2[0]) - k * _PIO2[1]) - k * _PIO2[2]
        z = r * r
        s1, s2, s3, s4 = _SIN_C
        c2, c3, c4, c5 = _COS_C
        ps = r + (r * z) * (s1 + z * (s2 + z * (s3 + z * s4)))
        pc = (_F(1.0) - _F(0.5) * z) + (z * z) * (c2 + z * (c3 + z * (c4 + z * c5)))
        q = np.where(k == k, k, _F(0.0)).astype(np.int64) & 3
        sn = np.where(q == 0, ps, np.where(q == 1, pc, np.where(q == 2, -ps, -pc)))
        cs = np.where(q == 0, pc, np.where(q == 1, -ps, np.where(q == 2, -pc, ps)))
        big = np.abs(x) > _SINCOS_LIMIT
        if np.any(big):
            xd = x.astype(np.float64)
            sn = np.where(big, np.sin(xd).astype(np.float32), sn)
            cs = np.where(big, np.cos(xd).astype(np.float32), cs)
    return sn.astype(np.float32), cs.astype(np.float32)


def exp_f32(x):
    """exp(x) in float32 by the device's reduction, polynomial and ldexp; the
    argument clamped to [-104, 88.5]."""
    x = np.asarray(x, dtype=np.float32)
    x = np.where(x < _F(-104.0), _F(-104.0), np.where(x > _F(88.5), _F(88.5), x)).astype(np.float32)
    with np.errstate(invalid="ignore"):
        n = np.rint(x * _LOG2E)
        r = (x - n * _LN2[0]) - n * _LN2[1]
        e1, e2, e3, e4, e5, e6, e7 = _EXP_C
        p = _F(1.0) + r * (e1 + r * (e2 + r * (e3 + r * (e4 + r * (e5 + r * (e6 + r * e7))))))
        ni = np.where(n == n, n, _F(0.0)).astype(np.int32)
    return np.ldexp(p, ni).astype(np.float32)


def _sincos(x):
    """(sin, cos) of the filament angles: the float32 routine, or float64."""
    x = np.asarray(x)
    if x.dtype == np.float32:
        return sincos_f32(x)
    return np.sin(x), np.cos(x)


def _cos(x):
    x = np.asarray(x)
    return sincos_f32(x)[1] if x.dtype == np.float32 else np.cos(x)


def _exp(x):
    x = np.asarray(x)
    return exp_f32(x) if x.dtype == np.float32 else np.exp(x)


def _tf2(fn, a, b):
    a = np.asarray(a)
    if a.dtype == np.float32:
        return fn(a.astype(np.float64), np.asarray(b).astype(np.float64)).astype(np.float32)
    return fn(a, b)


def _size_xy(grid_shape, dtype):
    """utils.py:21,32 -- jnp.flip(jnp.asarray((H, W))) -> (W, H) as floats."""
    h, w = grid_shape
    return np.array([w, h], dtype=dtype)


def displacement(grid_shape, x, y):
    """utils.py:18-23 -- minimum-image displacement mod(x - y + S/2, S) - S/2."""
    s = _size_xy(grid_shape, np.asarray(x).dtype)
    half = s / 2.0
    return np.mod(x - y + half, s) - half


def periodic_boundary(grid_shape, position):
    """utils.py:29-33 -- floor-mod of (x, y) by (W, H)."""
    return np.mod(position, _size_xy(grid_shape, position.dtype))


def index_from_position(position):
    """utils.py:25-27 -- (iy, ix) = flip(floor(pos)).T as integers."""
    return np.flip(np.floor(position).astype(np.int64), axis=-1).T


def fftfreq32(n, dtype):
    """jnp.fft.fftfreq(n) as JAX computes it: integer k cast to float, divided by n."""
    k = np.concatenate([np.arange(0, (n + 1) // 2), np.arange(-(n // 2), 0)]).astype(dtype)
    return k / np.asarray(n, dtype=dtype)


def gaussian_kernel_f(rows, cols, sigma, dtype):
    """utils.py:41-47 -- exp(-2 pi^2 sigma^2 (fr^2 + fc^2)) on the fftfreq grid."""
    r = fftfreq32(rows, dtype)
    c = fftfreq32(cols, dtype)
    r, c = np.meshgrid(r, c, indexing="ij")
    coef = -2 * np.pi ** 2 * sigma ** 2           # Python double, weak scalar
    return _tf(np.exp, coef * (r ** 2 + c ** 2))


def gaussian_filter_pocketfft(data, sigma):
    """utils.py:35-49 with numpy's FFT (used to pin the Stockham restatement)."""
    rows, cols = data.shape
    kf = gaussian_kernel_f(rows, cols, sigma, data.dtype)
    out = np.fft.ifft2(np.fft.fft2(data) * kf).real
    return out.astype(data.dtype)


def fft_radices(n):
    """Radix schedule of the device FFT: 4s, then 2, then odd factors <= 13."""
    out = []
    while n % 4 == 0:
        out.append(4); n //= 4
    while n % 2 == 0:
        out.append(2); n //= 2
    for p in range(3, 15, 2):
        while n % p == 0:
            out.append(p); n //= p
    if n != 1:
        raise ValueError("FFT length with a prime factor > 13")
    return out


def _twiddles(n, dtype):
    a = -2.0 * np.pi * np.arange(n, dtype=np.float64) / n
    return np.cos(a).astype(dtype), np.sin(a).astype(dtype)


def stockham_fft(re, im, inverse=False):
    """Unnormalised DFT along the last axis by the Stockham autosort network
    (radix 4 / 2 / 3 / generic), one rounding per real operation, in the order
    the device kernel performs them.  Batched over leading axes."""
    dtype = re.dtype
    n = re.shape[-1]
    twr, twi = _twiddles(n, dtype)
    if inverse:
        twi = -twi
    Ns = 1
    for R in fft_radices(n):
        nb = n // R
        tstride = n // (Ns * R)
        j = np.arange(nb)
        k = j % Ns
        vr, vi = [], []
        for r in range(R):
            xr = re[..., j + r * nb]
            xi = im[..., j + r * nb]
            if r > 0:
                m = k > 0
                wr = np.where(m, twr[(r * k * tstride) % n], 1).astype(dtype)
                wi = np.where(m, twi[(r * k * tstride) % n], 0).astype(dtype)
                yr = xr * wr - xi * wi
                yi = xr * wi + xi * wr
                xr = np.where(m, yr, xr)
                xi = np.where(m, yi, xi)
            vr.append(xr)
            vi.append(xi)
        if R == 2:
            Vr = [vr[0] + vr[1], vr[0] - vr[1]]
            Vi = [vi[0] + vi[1], vi[0] - vi[1]]
        elif R == 4:
            a0r, a0i = vr[0] + vr[2], vi[0] + vi[2]
            a1r, a1i = vr[0] - vr[2], vi[0] - vi[2]
            a2r, a2i = vr[1] + vr[3], vi[1] + vi[3]
            dr, di = vr[1] - vr[3], vi[1] - vi[3]
            a3r, a3i = (-di, dr) if inverse else (di, -dr)
            Vr = [a0r + a2r, a1r + a3r, a0r - a2r, a1r - a3r]
            Vi = [a0i + a2i, a1i + a3i, a0i - a2i, a1i - a3i]
        elif R == 3:
            s3 = dtype.type(0.866025403784438647)
            t1r, t1i = vr[1] + vr[2], vi[1] + vi[2]
            t2r, t2i = vr[1] - vr[2], vi[1] - vi[2]
            mr = vr[0] - dtype.type(0.5) * t1r
            mi = vi[0] - dtype.type(0.5) * t1i
            str_, sti = (-(s3 * t2i), s3 * t2r) if inverse else (s3 * t2i, -(s3 * t2r))
            Vr = [vr[0] + t1r, mr + str_, mr - str_]
            Vi = [vi[0] + t1i, mi + sti, mi - sti]
        else:
            Vr, Vi = [], []
            for q in range(R):
                ar, ai = vr[0], vi[0]
                for r in range(1, R):
                    idx = ((r * q) % R) * nb
                    wr, wi = twr[idx], twi[idx]
                    ar = ar + (vr[r] * wr - vi[r] * wi)
                    ai = ai + (vr[r] * wi + vi[r] * wr)
                Vr.append(ar)
                Vi.append(ai)
        dst = (j // Ns) * Ns * R + k
        ore = np.empty_like(re)
        oim = np.empty_like(im)
        for r in range(R):
            ore[..., dst + r * Ns] = Vr[r]
            oim[..., dst + r * Ns] = Vi[r]
        re, im = ore, oim
        Ns *= R
    return re, im


CONV_TAIL = 2.0 ** -27   # dropped kernel mass, relative to the kernel's total
CONV_TAPS_MAX = 129      # longest tap list of the direct filter (radius 64)


def gauss_taps(n, sigma):
    """The exact periodic 1-D kernel of utils.py:35-49's filter along an axis
    of length n: h[m] = IDFT(exp(-2 pi^2 sigma^2 f_k^2))[m] (the 2-D transfer
    function is the product of two such factors, so the filter is the
    circular convolution with h_y (x) h_x), in float64, k summed in order.
    Returns (lo, taps): taps[t] weights offset lo + t.  The radius R is the
    smallest with the mass of |h| beyond it at most CONV_TAIL of the total; a
    kernel that wraps (2 R + 1 >= n) keeps all n offsets once each.  None
    when more than CONV_TAPS_MAX taps are needed."""
    k = np.arange(n)
    f = np.where(k < (n + 1) // 2, k, k - n).astype(np.float64) / n
    T = np.exp((-2.0 * np.pi * np.pi * sigma * sigma) * (f * f))
    C = np.cos(2.0 * np.pi * np.arange(n, dtype=np.float64) / n)
    m = np.arange(n)
    h = np.zeros(n)
    for kk in range(n):
        h = h + T[kk] * C[(kk * m) % n]
    h = h / n
    a = np.abs(h)
    d = np.minimum(m, n - m)
    total = a.sum()
    R = 0
    while R < n // 2 and a[d > R].sum() > CONV_TAIL * total:
        R += 1
    if 2 * R + 1 >= n:
        lo, cnt = -((n - 1) // 2), n
    else:
        lo, cnt = -R, 2 * R + 1
    if cnt > CONV_TAPS_MAX:
        return None
    return lo, h[(lo + np.arange(cnt)) % n]


def gaussian_filter_direct(data, sigma, dtype=None):
    """The device's direct filter: circular convolution with the truncated
    kernels of ``gauss_taps``, rows (x) first, then columns (y), each output
    summed tap by tap from lo upward, rounded in ``dtype`` (default: the
    data's).  None when an axis needs more than CONV_TAPS_MAX taps."""
    dtype = np.dtype(dtype or data.dtype)
    H, W = data.shape
    tx, ty = gauss_taps(W, sigma), gauss_taps(H, sigma)
    if tx is None or ty is None:
        return None
    src = np.asarray(data, dtype)
    acc = np.zeros_like(src)
    for t, w in enumerate(tx[1].astype(dtype)):
        acc = acc + w * np.roll(src, -(tx[0] + t), axis=1)
    src, acc = acc, np.zeros_like(acc)
    for t, w in enumerate(ty[1].astype(dtype)):
        acc = acc + w * np.roll(src, -(ty[0] + t), axis=0)
    return acc


def gaussian_filter(data, sigma):
    """utils.py:35-49 -- Re(ifft2(fft2(data) * kernel_f)).

    float32 (the device's precision): the device's direct filter
    (``gaussian_filter_direct``) when both axes' kernels fit CONV_TAPS_MAX
    taps, else the device's FFT; float64: the FFT, exact to rounding."""
    if np.asarray(data).dtype == np.float32:
        out = gaussian_filter_direct(data, sigma)
        if out is not None:
            return out
    return gaussian_filter_fft(data, sigma)


def gaussian_filter_fft(data, sigma):
    """utils.py:35-49 -- Re(ifft2(fft2(data) * kernel_f)), the 2-D DFTs done as
    rows (x) forward, columns (y) forward, transfer function, columns inverse
    scaled by 1/H, rows inverse scaled by 1/W (the device's order)."""
    dtype = data.dtype
    rows, cols = data.shape
    kf = gaussian_kernel_f(rows, cols, sigma, dtype)
    re, im = stockham_fft(data.astype(dtype), np.zeros_like(data, dtype=dtype))
    re, im = stockham_fft(re.T.copy(), im.T.copy())           # columns, as rows of the transpose
    kt = kf.T
    re, im = re * kt, im * kt
    re, im = stockham_fft(re, im, inverse=True)
    sh = dtype.type(1.0 / rows)
    re, im = (re * sh).T.copy(), (im * sh).T.copy()
    re, im = stockham_fft(re, im, inverse=True)
    sw = dtype.type(1.0 / cols)
    return re * sw


# --------------------------------------------------------------------------
# Steric field (forces.py)
# --------------------------------------------------------------------------
def update_steric_potential(position, grid_shape, sigma, scatter_value):
    """forces.py:84-118 -- CIC deposit (corner-major, node-minor) + gaussian_filter."""
    dtype = position.dtype
    h, w = grid_shape
    steric = np.zeros((h, w), dtype=dtype)
    pos = periodic_boundary(grid_shape, position)
    x = pos[..., 0]
    y = pos[..., 1]
    fx = np.floor(x)
    fy = np.floor(y)
    x0 = np.mod(fx.astype(np.int32), w)
    y0 = np.mod(fy.astype(np.int32), h)
    x1 = np.mod(x0 + 1, w)
    y1 = np.mod(y0 + 1, h)
    dx = x - fx
    dy = y - fy
    w00 = (1.0 - dx) * (1.0 - dy) * scatter_value
    w10 = dx * (1.0 - dy) * scatter_value
    w01 = (1.0 - dx) * dy * scatter_value
    w11 = dx * dy * scatter_value
    np.add.at(steric, (y0, x0), w00)
    np.add.at(steric, (y0, x1), w10)
    np.add.at(steric, (y1, x0), w01)
    np.add.at(steric, (y1, x1), w11)
    return steric, gaussian_filter(steric, sigma)


def sample_bilinear_wrap(V, yy, xx):
    """jax.scipy.ndimage.map_coordinates(V, [yy, xx], order=1, mode="wrap").

    Linear weights (1 - t, t) per axis, corners visited in itertools.product
    order (y_lo,x_lo), (y_lo,x_hi), (y_hi,x_lo), (y_hi,x_hi), each weighted by
    wy * wx and summed left to right; indices wrap by floor-mod.
    """
    h, w = V.shape
    ly = np.floor(yy)
    wy_hi = yy - ly
    wy_lo = 1 - wy_hi
    iy = ly.astype(np.int32)
    lx = np.floor(xx)
    wx_hi = xx - lx
    wx_lo = 1 - wx_hi
    ix = lx.astype(np.int32)
    out = None
    for jy, wy in ((iy, wy_lo), (iy + 1, wy_hi)):
        for jx, wx in ((ix, wx_lo), (ix + 1, wx_hi)):
            c = (wy * wx) * V[np.mod(jy, h), np.mod(jx, w)]
            out = c if out is None else out + c
    return out


def force_interp(position, potential, fd_eps=0.5):
    """forces.py:20-38 -- -grad V by central differences of bilinear samples."""
    y = position[..., 1]
    x = position[..., 0]
    gx = (sample_bilinear_wrap(potential, y, x + fd_eps)
          - sample_bilinear_wrap(potential, y, x - fd_eps)) / (2.0 * fd_eps)
    gy = (sample_bilinear_wrap(potential, y + fd_eps, x)
          - sample_bilinear_wrap(potential, y - fd_eps, x)) / (2.0 * fd_eps)
    return -np.stack([gx, gy], axis=-1)


def steric_window_subtraction(pos, grid_shape, sigma, neighbor_skip, scatter_value):
    """forces.py:66-81 -- analytic (2n+1)-window contribution over the flat node
    index (the window crosses filament boundaries, as in the reference)."""
    dtype = pos.dtype
    n = pos.shape[0]
    offsets = np.arange(-neighbor_skip, neighbor_skip + 1)
    j_idx = np.arange(n)[:, None] + offsets[None, :]
    valid = ((j_idx >= 0) & (j_idx < n)).astype(dtype)
    j_safe = np.clip(j_idx, 0, n - 1)
    r = displacement(grid_shape, pos[:, None, :], pos[j_safe])
    d2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1]
    K = _exp(-d2 / (2.0 * sigma * sigma)) / (2.0 * np.pi * sigma * sigma)
    terms = (K * valid)[..., None] * r
    acc = np.zeros((n, 2), dtype=dtype)
    for k in range(terms.shape[1]):
        acc = acc + terms[:, k, :]
    return (scatter_value / (sigma * sigma)) * acc


def compute_field_steric_force_corrected(position, grid_shape, sigma, neighbor_skip,
                                         scatter_value=1.0):
    """forces.py:41-81 -- returns (density_grid, steric_field, f_steric)."""
    density, steric = update_steric_potential(position, grid_shape, sigma, scatter_value)
    pos = periodic_boundary(grid_shape, position)
    raw = force_interp(pos, steric)
    corr = steric_window_subtraction(pos, grid_shape, sigma, neighbor_skip, scatter_value)
    return density, steric, raw - corr


# --------------------------------------------------------------------------
# Rod solvers (solver/pbd.py, solver/stable_cosserat.py)
# --------------------------------------------------------------------------
def _edge_bending_degree(n_edges, bp, dtype):
    deg = np.zeros(n_edges, dtype=dtype)
    np.add.at(deg, bp[:, 0], 1.0)
    np.add.at(deg, bp[:, 1], 1.0)
    return np.maximum(deg, 1.0)


def _theta_bending_delta(theta, bp, rest, kb):
    """pbd.py:23-47 (without the final division) -- scatter of the wrapped
    hinge correction, e_in gets +c, e_out gets -c, in hinge order."""
    t_in = theta[bp[:, 0]]
    t_out = theta[bp[:, 1]]
    diff = t_out - t_in
    diff = np.mod(diff + np.pi, 2 * np.pi) - np.pi
    corr = (diff - rest) * kb * 0.5
    delta = np.zeros_like(theta)
    np.add.at(delta, bp[:, 0], corr)
    np.add.at(delta, bp[:, 1], -corr)
    return delta


def pbd_rod_constraint(s: State, p: Params):
    """pbd.py:14-91 -- one Jacobi iteration: theta pass (SOR 1.8), then the
    stretch/shear position pass.  Updates s.theta and s.position in place."""
    dtype = s.position.dtype
    src = s.pairs[:, 0]
    tgt = s.pairs[:, 1]
    bp = s.bending_pairs
    deg_e = _edge_bending_degree(s.theta.shape[0], bp, dtype)
    delta = _theta_bending_delta(s.theta, bp, s.bending_rest_angles, s.bending_stiffness)
    delta = delta / deg_e * 1.8
    s.theta = s.theta + delta

    n = s.position.shape[0]
    deg_n = np.zeros(n, dtype=dtype)
    np.add.at(deg_n, src, 1.0)
    np.add.at(deg_n, tgt, 1.0)
    deg_n = np.maximum(deg_n, 1.0)
    sn, cs = _sincos(s.theta)
    dirs = np.stack([cs, sn], axis=-1)
    target = dirs * s.rest_lengths[:, None]
    cur = displacement(s.grid_shape, s.position[tgt], s.position[src])
    length = np.sqrt(cur[:, 0:1] * cur[:, 0:1] + cur[:, 1:2] * cur[:, 1:2])
    cdir = cur / np.maximum(length, 1e-8)
    stretch = cdir * (length - s.rest_lengths[:, None])
    shear = (cur - target) - stretch
    forces = (stretch + shear * p.stiffness_shear) * p.stiffness_stretch * 0.5
    disp = np.zeros_like(s.position)
    np.add.at(disp, src, forces)
    np.add.at(disp, tgt, -forces)
    disp = disp / deg_n[:, None]
    s.position = s.position + disp


def stable_cosserat_constraint(s: State, p: Params):
    """stable_cosserat.py:13-119 -- theta bending pass + 0.05 shear pull, then the
    implicit per-vertex update with inertia w = 1/config.dt^2 toward y = debug_vector."""
    dtype = s.position.dtype
    x = s.position
    y = s.debug_vector
    src = s.pairs[:, 0]
    tgt = s.pairs[:, 1]
    l = s.rest_lengths
    bp = s.bending_pairs
    deg_e = _edge_bending_degree(s.theta.shape[0], bp, dtype)
    delta = _theta_bending_delta(s.theta, bp, s.bending_rest_angles, s.bending_stiffness)
    delta = delta / deg_e
    theta_new = s.theta + delta
    seg = displacement(s.grid_shape, x[tgt], x[src])
    alpha = _tf2(np.arctan2, seg[:, 1], seg[:, 0])
    pull = np.mod(alpha - theta_new + np.pi, 2 * np.pi) - np.pi
    theta_new = theta_new + pull * 0.05
    s.theta = theta_new

    k_s = p.stiffness_stretch
    w = 1.0 / (p.config_dt ** 2)
    sn, cs = _sincos(theta_new)
    d = np.stack([cs, sn], axis=-1)
    r = displacement(s.grid_shape, x[tgt], x[src])
    c = r / l[:, None] - d
    f_edge = (k_s / l)[:, None] * c
    h_edge = k_s / (l ** 2)
    f_node = w * (y - x)
    h_node = np.full(x.shape[0], w, dtype=dtype)
    np.add.at(f_node, src, f_edge)
    np.add.at(f_node, tgt, -f_edge)
    np.add.at(h_node, src, h_edge)
    np.add.at(h_node, tgt, h_edge)
    s.position = x + f_node / h_node[:, None]


# --------------------------------------------------------------------------
# Immersed boundary + LBM (solver/fluid.py)
# --------------------------------------------------------------------------
def kernel_1d_peskin(r, h):
    """fluid.py:31-38 -- 0.25 (1 + cos(pi |r|/h / 2)) for |r|/h <= 2."""
    r_abs = np.abs(r) / h
    return np.where(r_abs <= 2.0, 0.25 * (1.0 + _cos(np.pi * r_abs / 2.0)), 0.0).astype(r.dtype)


def ib_stencil(grid_pos, h, w, kernel_size):
    """fluid.py:40-87 -- per-node flat int32 cell indices and normalised weights,
    shape (N, K*K).  Tap t has dx = t // K (outer) and dy = t % K (inner): the
    first output of meshgrid(offsets, offsets) is assigned to dy (fluid.py:54)."""
    dtype = grid_pos.dtype
    k = kernel_size
    half = (k - 1) // 2
    half_width = k / 2.0
    x = grid_pos[:, 0:1]
    y = grid_pos[:, 1:2]
    start_x = np.floor(x) - half
    start_y = np.floor(y) - half
    offsets = np.arange(float(k)).astype(dtype)
    dy, dx = np.meshgrid(offsets, offsets)
    nx = start_x + dx.ravel()[None, :]
    ny = start_y + dy.ravel()[None, :]
    wx = kernel_1d_peskin(nx - x, half_width)
    wy = kernel_1d_peskin(ny - y, half_width)
    weights = wx * wy
    total = np.zeros_like(weights[:, 0])
    for t in range(k * k):
        total = total + weights[:, t]
    weights = weights / (total + EPSIL)[:, None]
    idx_x = np.mod(nx.astype(np.int32), w)
    idx_y = np.mod(ny.astype(np.int32), h)
    flat = (idx_y * w + idx_x).astype(np.int32)
    return flat, weights


def ib_interpolate(flat, weights, field2):
    """fluid.py:90-97 -- per node sum_t w_t * field[:, flat_t], left fold over taps."""
    vflat = field2.reshape(2, -1)
    acc = np.zeros((flat.shape[0], 2), dtype=field2.dtype)
    for t in range(flat.shape[1]):
        acc = acc + vflat[:, flat[:, t]].T * weights[:, t:t + 1]
    return acc


def ib_spread(flat, weights, force, h, w):
    """fluid.py:109-120 -- in-order scatter of force * weight (node-major, tap-minor)."""
    nst = flat.shape[1]
    idx = flat.reshape(-1)
    wexp = weights.reshape(-1)
    out = np.zeros((2, h * w), dtype=force.dtype)
    np.add.at(out[0], idx, np.repeat(force[:, 0], nst) * wexp)
    np.add.at(out[1], idx, np.repeat(force[:, 1], nst) * wexp)
    return out.reshape(2, h, w)


def ib_kernel_weight(flat, weights, h, w):
    """fluid.py:123-133 -- in-order scatter of the weights."""
    out = np.zeros(h * w, dtype=weights.dtype)
    np.add.at(out, flat.reshape(-1), weights.reshape(-1))
    return out.reshape(h, w)


def macroscopic(f):
    """fluid.py:136-142 -- rho = max(sum f, 0.1), u = (sum e f) / rho; left folds."""
    rho = f[0]
    for i in range(1, 9):
        rho = rho + f[i]
    rho = np.maximum(rho, 0.1)
    mom = np.zeros((2,) + f.shape[1:], dtype=f.dtype)
    for i in range(9):
        for a in range(2):
            e = int(LBM_VELOCITIES[i, a])
            if e != 0:
                mom[a] = mom[a] + e * f[i]
    return rho, mom / rho


def equilibrium(rho, u):
    """fluid.py:145-151 -- rho w_i (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2)."""
    dtype = rho.dtype
    usq = u[0] * u[0] + u[1] * u[1]
    out = np.empty((9,) + rho.shape, dtype=dtype)
    for i in range(9):
        eu = _edot(i, u)
        out[i] = rho * dtype.type(LBM_WEIGHTS[i]) * (1.0 + 3.0 * eu + 4.5 * (eu * eu) - 1.5 * usq)
    return out


def _edot(i, v):
    """e_i . v for a (2, ...) field; products with 0/+-1 are exact, two-term sum."""
    ex, ey = (int(c) for c in LBM_VELOCITIES[i])
    return ex * v[0] + ey * v[1]


def _clamp_mag(u, umax):
    """fluid.py:272-274 -- u * min(1, umax / (sqrt(|u|^2 + eps) + eps))."""
    mag = np.sqrt(u[0] * u[0] + u[1] * u[1] + EPSIL)
    return u * np.minimum(1.0, umax / (mag + EPSIL))[None]


def trt_omegas(viscosity, lambda_trt, dtype):
    """fluid.py:155-161,342-344 -- omega+ from nu (double, then clipped in the
    array precision); omega- from the magic parameter in the array precision."""
    op = np.clip(np.asarray(1.0 / (3.0 * viscosity + 0.5), dtype=dtype), 0.0, 1.9).astype(dtype)
    tpo = 1.0 / np.maximum(op, EPSIL) - 0.5
    om = 1.0 / (lambda_trt / np.maximum(tpo, EPSIL) + 0.5)
    return dtype.type(op), dtype.type(om)


def collide_trt_guo(f, rho, u, force, omega_plus, omega_minus):
    """fluid.py:164-199,346-350 -- TRT collision about feq(rho, u*) plus the TRT
    Guo source; u* = u + 0.5 F / rho; floor at EPSIL."""
    u_star = u + 0.5 * force / rho
    feq = equilibrium(rho, u_star)
    dtype = f.dtype
    raw = np.empty_like(f)
    udf = u_star[0] * force[0] + u_star[1] * force[1]
    for i in range(9):
        eu = _edot(i, u_star)
        ef = _edot(i, force)
        raw[i] = dtype.type(LBM_WEIGHTS[i]) * (3.0 * (ef - udf) + 9.0 * eu * ef)
    out = np.empty_like(f)
    cp = 1.0 - 0.5 * omega_plus
    cm = 1.0 - 0.5 * omega_minus
    for i in range(9):
        o = OPP_INDICES[i]
        coll = (f[i] - omega_plus * (0.5 * (f[i] + f[o]) - 0.5 * (feq[i] + feq[o]))
                - omega_minus * (0.5 * (f[i] - f[o]) - 0.5 * (feq[i] - feq[o])))
        src = cp * (0.5 * (raw[i] + raw[o])) + cm * (0.5 * (raw[i] - raw[o]))
        out[i] = coll + src
    return np.maximum(out, EPSIL), u_star


def stream(f_post):
    """fluid.py:352-357 -- f_new[i][y, x] = f_post[i][y - e_iy, x - e_ix] (periodic)."""
    out = np.empty_like(f_post)
    for i in range(9):
        ex, ey = LBM_VELOCITIES[i]
        out[i] = np.roll(f_post[i], shift=(int(ey), int(ex)), axis=(0, 1))
    return np.maximum(out, EPSIL)


@dataclass
class IBTrace:
    """Intermediates of one immersed_boundary_interaction call (for tests)."""
    flat: np.ndarray = None
    weights: np.ndarray = None
    kernel_weight: np.ndarray = None
    u_corrected: np.ndarray = None
    force: np.ndarray = None
    rho: np.ndarray = None
    velocity: np.ndarray = None
    node_force: np.ndarray = None


def synthetic_ib_points(position, pairs, grid_shape, grid_pos, width, mean_scale):
    """fluid.py:246-263 -- the two synthetic IB points of every node, at
    +/- width * mean_scale along the chain normal: rows 0..N-1 the + points,
    N..2N-1 the - points.  The tangent is the sum of the node's incident edge
    vectors (tgt - src, minimum image), added in the reference's scatter
    order: every edge at its source in edge order, then at its target in edge
    order; the normal is the unit tangent turned by +90 degrees."""
    dtype = grid_pos.dtype
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    src, tgt = pairs[:, 0], pairs[:, 1]
    ev = displacement(grid_shape, position[tgt], position[src])
    traw = np.zeros_like(position)
    np.add.at(traw, src, ev)  # sequential in index order
    np.add.at(traw, tgt, ev)
    nrm = np.sqrt((traw[:, 0:1] * traw[:, 0:1] + traw[:, 1:2] * traw[:, 1:2]) + EPSIL)
    t = traw / nrm
    perp = np.concatenate([-t[:, 1:2], t[:, 0:1]], axis=1)
    off = (np.asarray(width, dtype) * perp) * np.asarray(mean_scale, dtype)
    return np.concatenate([grid_pos + off, grid_pos - off], axis=0)


def immersed_boundary_interaction(position, velocity, grid_shape, f_grid, dt, p: Params,
                                  trace: Optional[IBTrace] = None, pairs=None):
    """fluid.py:202-386 -- IB multi-direct forcing + TRT/Guo collide + stream + drag.

    With synthetic_node_width > 0 the IB runs on the 2N synthetic points of
    synthetic_ib_points (``pairs`` is then required) and each node's drag is
    the mean of its two points' forces (fluid.py:364-368).
    Returns (new node velocity, fluid_velocity = u*, f_new).
    """
    dtype = f_grid.dtype
    h, w = f_grid.shape[1], f_grid.shape[2]
    scale_x = w / grid_shape[0]
    scale_y = h / grid_shape[1]
    scale = np.array([scale_x, scale_y], dtype=dtype)
    mean_scale = 0.5 * (scale_x + scale_y)
    umax = p.max_fluid_velocity * mean_scale
    ibm_vel = velocity * scale * dt
    grid_pos = position * scale
    n_nodes = position.shape[0]
    synthetic = p.synthetic_node_width > 0.0
    if synthetic:
        if pairs is None:
            raise ValueError("synthetic_node_width > 0 needs the edge pairs")
        grid_pos = synthetic_ib_points(position, pairs, grid_shape, grid_pos,
                                       p.synthetic_node_width, mean_scale)
        ibm_vel = np.concatenate([ibm_vel, ibm_vel], axis=0)

    f = np.maximum(f_grid, EPSIL)
    rho, u = macroscopic(f)
    u = _clamp_mag(u, umax)

    flat, wts = ib_stencil(grid_pos, h, w, p.ibm_kernel_size)
    kw = ib_kernel_weight(flat, wts, h, w)
    inv_kw = 1.0 / np.maximum(kw[None], 1.0)

    u_corr = u
    for _ in range(p.ibm_iterations):
        un = ib_interpolate(flat, wts, u_corr)
        c = (ibm_vel - un) * p.ibm_relaxation
        mag = np.sqrt(c[:, 0:1] * c[:, 0:1] + c[:, 1:2] * c[:, 1:2] + EPSIL)
        c = c * np.minimum(1.0, umax / (mag + EPSIL))
        delta = ib_spread(flat, wts, c, h, w) * inv_kw
        u_corr = u_corr + delta

    u_corr = _clamp_mag(u_corr, umax)
    force = (u_corr - u) * rho
    if p.enable_force_clamp:
        force = _clamp_mag(force, umax)

    op, om = trt_omegas(p.viscosity, p.lambda_trt, dtype)
    f_post, u_star = collide_trt_guo(f, rho, u, force, op, om)
    f_new = stream(f_post)

    fn = ib_interpolate(flat, wts, force)
    if synthetic:
        fn = (fn[:n_nodes] + fn[n_nodes:]) * dtype.type(0.5)
    dv = (fn / p.node_mass) / (scale * dt)
    if trace is not None:
        trace.flat, trace.weights, trace.kernel_weight = flat, wts, kw
        trace.u_corrected, trace.force, trace.rho, trace.velocity = u_corr, force, rho, u
        trace.node_force = fn
    return velocity - dv, u_star, f_new


# --------------------------------------------------------------------------
# The step (simulate.py)
# --------------------------------------------------------------------------
def step(s: State, dt: float, p: Params, ib_position=None, ib_trace=None) -> State:
    """simulate.py:37-100 -- one coupled step; returns a new State.

    For checking gradients, which the reference takes with the IB positions
    stop_gradient'ed (fluid.py:239): ``ib_trace`` (a list) receives the
    positions the IB stencil is built from, and ``ib_position`` replaces
    them, so that a finite difference with the recorded positions held fixed
    is the derivative the reference's jax.grad computes."""
    s = s.copy()
    s.debug_vector = np.zeros_like(s.position)
    prev = s.position
    s.velocity = s.velocity * p.damping
    pred = prev + s.velocity * dt
    if p.enable_steric:
        _, s.steric, fst = compute_field_steric_force_corrected(
            pred, s.grid_shape, p.steric_sigma, p.steric_neighbor_skip, p.steric_scatter_value)
        pred = pred + fst * p.steric_strength * dt
    s.position = pred
    # solver_type only sets the inertial target (simulate.py:71-72); the
    # constraint run is bending_constraint (simulate.py:74-77)
    if p.solver_type == "cosserat":
        s.debug_vector = pred
    bc = p.bending_constraint or p.solver_type
    fn = stable_cosserat_constraint if bc == "cosserat" else pbd_rod_constraint
    for _ in range(p.cycles_per_step):
        fn(s, p)
    s.velocity = (s.position - prev) / dt
    if p.enable_fluid:
        if ib_trace is not None:
            ib_trace.append(s.position.copy())
        s.velocity, s.fluid_velocity, s.f_grid = immersed_boundary_interaction(
            s.position if ib_position is None else ib_position, s.velocity, s.grid_shape,
            s.f_grid, dt, p, pairs=s.pairs)
    s.position = periodic_boundary(s.grid_shape, s.position)
    return s


def resize_weights(n_in, n_out, dtype=np.float64):
    """The (n_in, n_out) weight matrix of jax.image.resize(..., method="linear")
    along one axis (the published triangle-kernel scale-and-translate with
    antialiasing that simulate.py:24-28 calls; jax 0.8.1, not vendored): the
    sample position of output o is (o + 0.5) / scale - 0.5 with scale =
    n_out / n_in; input i weighs max(0, 1 - |sample - i| / max(1 / scale, 1));
    each output's weights are divided by their sum (when it exceeds 1000 eps)
    and zeroed when the sample lies outside [-0.5, n_in - 0.5]."""
    dt = np.dtype(dtype).type
    scale = n_out / n_in
    inv = 1.0 / scale
    ks = max(inv, 1.0)
    sf = (np.arange(n_out, dtype=dt) + dt(0.5)) * dt(inv) - dt(0.5)
    x = np.abs(sf[None, :] - np.arange(n_in, dtype=dt)[:, None]) / dt(ks)
    w = np.maximum(dt(0.0), dt(1.0) - np.abs(x)).astype(dt)
    tot = np.zeros((1, n_out), dt)
    for i in range(n_in):  # in input order
        tot = tot + w[i:i + 1]
    w = np.where(np.abs(tot) > dt(1000.0) * EPSIL, w / np.where(tot != 0, tot, dt(1)), dt(0.0))
    inside = (sf >= dt(-0.5)) & (sf <= dt(n_in - 0.5))
    return np.where(inside[None, :], w, dt(0.0)).astype(dt)


def resize_linear(u, shape):
    """simulate.py:24-28 -- jax.image.resize(u, shape, method="linear") of a
    (C, H, W) array to (C, H', W'), separable: rows, then columns."""
    c, h, w = u.shape
    _, h2, w2 = shape
    out = np.asarray(u)
    if (h2, w2) == (h, w):
        return out.copy()
    wy = resize_weights(h, h2, np.float64) if h2 != h else np.eye(h)
    wx = resize_weights(w, w2, np.float64) if w2 != w else np.eye(w)
    return np.einsum("chw,hy,wx->cyx", out.astype(np.float64), wy, wx)


def simulate(s: State, dt: float, num_steps: int, p: Params, actuation=None):
    """simulate.py:15-35 -- returns (final state, per-step positions, per-step u*).

    ``actuation(t, state) -> bending_rest_angles`` (optional) is applied before
    each step, the way the gym/worm callers set rest angles (gym/base.py:169-173).
    """
    pos_ts, vel_ts = [], []
    for t in range(num_steps):
        if actuation is not None:
            s = replace(s, bending_rest_angles=actuation(t, s))
        s = step(s, dt, p)
        pos_ts.append(s.position)
        vel_ts.append(s.fluid_velocity)
    return s, np.stack(pos_ts), np.stack(vel_ts)


# --------------------------------------------------------------------------
# Topology and initial state (graph.py, env.py)
# --------------------------------------------------------------------------
def make_edges(num_nodes, graph="loop", num_instances=1):
    """graph.py:156-259 -- concatenated per-instance edge lists + departure angles."""
    per = num_nodes // num_instances
    rem = num_nodes % num_instances
    if per < 1:
        raise ValueError("num_nodes must be at least num_instances")
    edges, deps = [], []
    start = 0
    for i in range(num_instances):
        size = per + (1 if i < rem else 0)
        if graph == "loop":
            src = np.arange(size) + start
            e = np.stack([src, np.roll(src, -1)], axis=1)
            d = np.zeros(size)
        elif graph == "line":
            src = np.arange(size - 1) + start
            e = np.stack([src, src + 1], axis=1)
            d = np.zeros(size - 1)
        elif graph == "tree":
            trunk = int(size * 0.6)
            remn = size - trunk
            b1 = remn // 2
            b2 = remn - b1
            t0, b1s = start, start + trunk
            b2s = b1s + b1
            junction = t0 + trunk // 2
            el, dl = [], []
            for k in range(trunk - 1):
                el.append((t0 + k, t0 + k + 1)); dl.append(0.0)
            el.append((junction, b1s)); dl.append(np.pi / 2.0)
            for k in range(b1 - 1):
                el.append((b1s + k, b1s + k + 1)); dl.append(np.pi / 2.0)
            el.append((junction, b2s)); dl.append(-np.pi / 2.0)
            for k in range(b2 - 1):
                el.append((b2s + k, b2s + k + 1)); dl.append(-np.pi / 2.0)
            e = np.array(el)
            d = np.array(dl)
        else:
            raise ValueError(f"Unknown graph type: {graph}")
        edges.append(e)
        deps.append(d)
        start += size
    return (np.concatenate(edges).astype(np.int32),
            np.concatenate(deps).astype(np.float32))


def compute_bending_pairs(edge_pairs, departure_angles, num_nodes):
    """graph.py:347-401 -- per node (ascending), every incoming x outgoing edge
    pair in edge order; rest angle = dep[e_out] - dep[e_in]."""
    ep = np.asarray(edge_pairs)
    dep = np.asarray(departure_angles)
    inc = [[] for _ in range(num_nodes)]
    out = [[] for _ in range(num_nodes)]
    for e in range(ep.shape[0]):
        out[int(ep[e, 0])].append(e)
        inc[int(ep[e, 1])].append(e)
    bp, rest = [], []
    for v in range(num_nodes):
        for ei in inc[v]:
            for eo in out[v]:
                bp.append((ei, eo))
                rest.append(float(dep[eo] - dep[ei]))
    if not bp:
        return np.zeros((0, 2), np.int32), np.zeros(0, np.float32)
    return np.array(bp, np.int32), np.array(rest, np.float32)


def initialize_f_from_velocity(velocity_field, density=1.0):
    """graph.py:533-540 -- feq(density, u) with float32 lattice weights."""
    dtype = velocity_field.dtype
    u = velocity_field
    usq = u[0] * u[0] + u[1] * u[1]
    out = np.empty((9,) + u.shape[1:], dtype=dtype)
    for i in range(9):
        eu = _edot(i, u)
        out[i] = dtype.type(density * LBM_WEIGHTS[i]) * (1.0 + 3.0 * eu + 4.5 * (eu * eu) - 1.5 * usq)
    return out


# --------------------------------------------------------------------------
# Swimmer fitness (the caller of the step)
# --------------------------------------------------------------------------
COM_LANES = 256  # the device reduction width


def _lane_tree_mean(d):
    """Mean of d (n, 2) in the device's order: lane i folds rows i, i + 256,
    ... in order from 0, then the 256 lane sums are halved pairwise
    (p[i] += p[i + s], s = 128 .. 1), then divided by n."""
    n = d.shape[0]
    dt = d.dtype
    part = np.zeros((COM_LANES, 2), dt)
    for k in range(0, n, COM_LANES):
        blk = d[k:k + COM_LANES]
        part[:blk.shape[0]] = part[:blk.shape[0]] + blk
    s = COM_LANES // 2
    while s > 0:
        part[:s] = part[:s] + part[s:2 * s]
        s //= 2
    return part[0] / dt.type(n)


def group_com(position, grid_shape, offsets=None, periodic=True):
    """Centre of each node group [off[g], off[g+1]).

    periodic=True: swim_base.py:167-177 ``_periodic_com`` -- displacements of
    every node from the group's first node (utils.py:18-23), their mean, added
    to the first node and wrapped (utils.py:29-33).  periodic=False: the plain
    mean of 01_es_locomotion.ipynb ``rollout_fitness``.  The mean's summation
    order is the device's (``_lane_tree_mean``); jnp.mean's XLA order is not
    reproduced (it differs in the last bits only)."""
    pos = np.asarray(position)
    off = [0, pos.shape[0]] if offsets is None else list(offsets)
    out = np.zeros((len(off) - 1, 2), pos.dtype)
    for g in range(len(off) - 1):
        p = pos[off[g]:off[g + 1]]
        if periodic:
            ref = p[0]
            m = _lane_tree_mean(displacement(grid_shape, p, ref))
            out[g] = periodic_boundary(grid_shape, ref + m)
        else:
            out[g] = _lane_tree_mean(p)
    return out


def swim_track(positions, grid_shape, offsets=None, start_step=0):
    """swim_base.py:179-209 over a trajectory: positions[0] initialises
    (``_init_displacement``), each later frame is one ``_update_displacement``.
    Returns the cumulative displacement (G, 2) after every update."""
    prev = group_com(positions[0], grid_shape, offsets)
    cum = np.zeros_like(prev)
    hist = []
    for k, pos in enumerate(positions[1:], start=1):
        com = group_com(pos, grid_shape, offsets)
        delta = displacement(grid_shape, com, prev)
        prev = com
        if k > start_step:
            cum = cum + delta
        hist.append(cum.copy())
    return np.asarray(hist)
