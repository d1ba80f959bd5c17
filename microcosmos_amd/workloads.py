"""Seeded synthetic workloads for the coupled fluid-filament step (host, numpy).

The workloads are plain initial states (positions, topology, lattice) plus an
actuation schedule for the bending rest angles.  They are produced with numpy
``default_rng`` seeds so the device path and the CPU restatement receive
bit-identical inputs.

Topology follows graph.py:156-401 (``make_edges`` line instances concatenated,
``compute_bending_pairs``); the lattice starts at feq(1, u0) as in
graph.py:533-540.  The actuation mirrors how callers drive the step by
rewriting ``bending_rest_angles`` before each call (gym/base.py:169-173,
swim_sweep/swimmers/worm.py:60-74):

    rest_b(t) = amp_b * sin(omega_b * t + phase_b)      (evaluated in float64)
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import topology

LBM_E = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                  [1, 1], [-1, 1], [-1, -1], [1, -1]])
LBM_W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])


@dataclass
class Workload:
    """Initial state + solver overrides of one synthetic configuration."""
    name: str
    grid_shape: tuple
    position: np.ndarray
    velocity: np.ndarray
    pairs: np.ndarray
    theta: np.ndarray
    rest_lengths: np.ndarray
    bending_pairs: np.ndarray
    bending_rest_angles: np.ndarray
    bending_stiffness: np.ndarray
    f_grid: np.ndarray
    fluid_velocity: np.ndarray
    steric: np.ndarray
    dt: float
    config_overrides: dict
    act_amp: np.ndarray
    act_omega: np.ndarray
    act_phase: np.ndarray
    steps: int = 1000
    color_bend: Optional[np.ndarray] = None

    @property
    def num_nodes(self):
        return self.position.shape[0]

    def rest_angles_at(self, t: int) -> np.ndarray:
        """Bending rest angles applied before step ``t`` (float64 evaluation)."""
        arg = self.act_omega.astype(np.float64) * float(t) + self.act_phase.astype(np.float64)
        return (self.act_amp.astype(np.float64) * np.sin(arg)).astype(np.float32)


def feq_init(u0: np.ndarray) -> np.ndarray:
    """graph.py:533-540 -- feq(rho = 1, u0) in float32 with float32 weights."""
    u = u0.astype(np.float32)
    usq = u[0] * u[0] + u[1] * u[1]
    out = np.empty((9,) + u.shape[1:], np.float32)
    for i in range(9):
        ex, ey = int(LBM_E[i, 0]), int(LBM_E[i, 1])
        eu = ex * u[0] + ey * u[1]
        out[i] = np.float32(1.0 * LBM_W[i]) * (1.0 + 3.0 * eu + 4.5 * (eu * eu) - 1.5 * usq)
    return out


def _min_image(grid_shape, a, b):
    h, w = grid_shape
    s = np.array([w, h], np.float32)
    half = s / 2.0
    return np.mod(a - b + half, s) - half


def filaments_workload(name, grid, n_fil, n_nodes, seed, *, straight_center=False,
                       jitter=0.05, spacing=1.0, amp=0.3, period=200.0, kb=0.5,
                       u0_amp=1e-3, steps=1000, overrides=None, node_counts=None,
                       amps=None, periods=None):
    """``n_fil`` line filaments of ``n_nodes`` nodes on a periodic ``grid``.

    straight_center: one filament centred horizontally (the 128^2 plain case);
    otherwise start points are uniform on the grid, headings uniform in
    [0, 2 pi), actuation phases uniform in [0, 2 pi).  amps / periods: per
    filament actuation amplitude and period (default: amp, period for all).
    """
    rng = np.random.default_rng(seed)
    h, w = grid
    u0 = rng.uniform(-u0_amp, u0_amp, size=(2, h, w)).astype(np.float32)
    counts = list(node_counts) if node_counts is not None else [n_nodes] * n_fil
    pos_list = []
    phases = []
    for f, n in enumerate(counts):
        i = np.arange(n, dtype=np.float64)
        if straight_center:
            x = w / 2.0 + (i - (n - 1) / 2.0) * spacing
            y = np.full(n, h / 2.0)
            phi = 0.0
        else:
            x0, y0 = rng.uniform(0, w), rng.uniform(0, h)
            head = rng.uniform(0, 2 * np.pi)
            x = x0 + i * spacing * np.cos(head)
            y = y0 + i * spacing * np.sin(head)
            phi = rng.uniform(0, 2 * np.pi)
        p = np.stack([x, y], axis=1) + rng.normal(0.0, jitter, size=(n, 2))
        p = np.mod(p, np.array([w, h], np.float64))
        pos_list.append(p.astype(np.float32))
        phases.append(phi)
    position = np.concatenate(pos_list).astype(np.float32)
    pairs, bp, rest0 = topology.line_instances(counts)
    d = _min_image(grid, position[pairs[:, 1]], position[pairs[:, 0]])
    theta = np.arctan2(d[:, 1].astype(np.float64), d[:, 0].astype(np.float64)).astype(np.float32)
    # actuation: travelling wave along each filament's hinges
    amp_b, om_b, ph_b = [], [], []
    start_b = 0
    for f, n in enumerate(counts):
        nb = max(n - 2, 0)
        b = np.arange(nb, dtype=np.float64)
        a_f = amp if amps is None else float(amps[f])
        t_f = period if periods is None else float(periods[f])
        amp_b.append(np.full(nb, a_f))
        om_b.append(np.full(nb, 2 * np.pi / t_f))
        ph_b.append(phases[f] - 2 * np.pi * b / max(nb, 1))
        start_b += nb
    act_amp = np.concatenate(amp_b).astype(np.float32)
    act_omega = np.concatenate(om_b).astype(np.float32)
    act_phase = np.concatenate(ph_b).astype(np.float32)
    wl = Workload(
        name=name, grid_shape=(h, w), position=position,
        velocity=np.zeros_like(position), pairs=pairs, theta=theta,
        rest_lengths=np.full(pairs.shape[0], spacing, np.float32),
        bending_pairs=bp, bending_rest_angles=rest0.astype(np.float32),
        bending_stiffness=np.full(bp.shape[0], kb, np.float32),
        f_grid=feq_init(u0), fluid_velocity=np.zeros((2, h, w), np.float32),
        steric=np.zeros((h, w), np.float32), dt=1.0,
        config_overrides=dict(overrides or {}), act_amp=act_amp,
        act_omega=act_omega, act_phase=act_phase, steps=steps,
        color_bend=(np.arange(position.shape[0]) % 3).astype(np.int32))
    return wl


def _lbm_overrides(tau_plus, cycles):
    return dict(viscosity=(tau_plus - 0.5) / 3.0, lambda_trt=0.1875,
                cycles_per_step=cycles, ibm_kernel_size=4, enable_steric=True,
                steric_sigma=2.0)


def plain_128(seed=0, steps=1000):
    """128^2 periodic D2Q9 TRT (tau+ 0.8, Lambda 3/16), one 32-node filament at
    the centre, IB K=4, 4 PBD cycles, steric sigma 2, dt 1."""
    return filaments_workload("plain-128x128-1x32", (128, 128), 1, 32, seed,
                              straight_center=True, steps=steps,
                              overrides=_lbm_overrides(0.8, 4))


def plain_1024(seed=1, steps=1000):
    """1024^2 periodic D2Q9 TRT (tau+ 0.6), 256 filaments x 64 nodes uniformly
    placed and headed, IB K=4, 8 PBD cycles, steric sigma 2."""
    return filaments_workload("plain-1024x1024-256x64", (1024, 1024), 256, 64, seed,
                              steps=steps, overrides=_lbm_overrides(0.6, 8))


def small(grid=(48, 48), n_fil=2, n_nodes=16, seed=7, cycles=4, tau_plus=0.8, steps=100,
          straight_center=False, **kw):
    """Small seeded case for parity tests."""
    return filaments_workload(f"small-{grid[0]}x{grid[1]}-{n_fil}x{n_nodes}", grid, n_fil,
                              n_nodes, seed, straight_center=straight_center, steps=steps,
                              overrides=_lbm_overrides(tau_plus, cycles), **kw)


def es_population(n=512, seed=2, sides=(128, 192, 256), max_filaments=16, min_nodes=16,
                  max_nodes=64, steps=500):
    """Batched ES population: ``n`` independent instances, each with a grid
    side drawn from ``sides``, 1..max_filaments filaments of
    min_nodes..max_nodes nodes, per-filament actuation amplitude U(0, 0.5) and
    period U(100, 400); fluid, IB and steric settings of the 128^2 plain case
    (tau+ 0.8, Lambda 3/16, K = 4, 4 PBD cycles, sigma 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(rng.choice(sides))
        nf = int(rng.integers(1, max_filaments + 1))
        counts = rng.integers(min_nodes, max_nodes + 1, size=nf).tolist()
        amps = rng.uniform(0.0, 0.5, size=nf)
        periods = rng.uniform(100.0, 400.0, size=nf)
        out.append(filaments_workload(
            f"es-{i}-{side}x{side}-{nf}f", (side, side), nf, 0, int(rng.integers(1 << 30)),
            node_counts=counts, amps=amps, periods=periods, steps=steps,
            overrides=_lbm_overrides(0.8, 4)))
    return out


@dataclass
class ServeItem:
    """One request of a serving trace: arrival time (s), workload, steps and
    whether it is a gradient request (then ``target`` is its folding target
    curve and the workload a folding workload)."""
    arrival: float
    workload: Workload
    steps: int
    gradient: bool
    target: Optional[np.ndarray] = None


def serving_trace(n=48, seed=4, rate=200.0, sides=(128, 256, 512, 1024), max_filaments=256,
                  min_nodes=16, max_nodes=64, min_steps=200, max_steps=2000,
                  gradient_fraction=0.2):
    """Poisson(rate) arrivals of ``n`` requests.  A fraction
    ``gradient_fraction`` are gradient requests: the folding gradient of
    ``folding()`` (256^2, 8 filaments x 48 nodes, 500 steps) with a seeded
    target.  The others are rollouts: grid side uniform over ``sides``,
    1..max_filaments filaments of min_nodes..max_nodes nodes, rollouts of
    min_steps..max_steps steps, per-filament actuation amplitudes U(0, 0.5)
    and periods U(100, 400), fluid, IB and steric settings of the 128^2 plain
    case."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        side = int(rng.choice(sides))
        nf = int(rng.integers(1, max_filaments + 1))
        counts = rng.integers(min_nodes, max_nodes + 1, size=nf).tolist()
        amps = rng.uniform(0.0, 0.5, size=nf)
        periods = rng.uniform(100.0, 400.0, size=nf)
        steps = int(rng.integers(min_steps, max_steps + 1))
        grad = bool(rng.uniform() < gradient_fraction)
        wseed = int(rng.integers(1 << 30))
        if grad:
            wl, tgt = folding(seed=wseed % 100000)
            out.append(ServeItem(arrival=t, workload=wl, steps=wl.steps, gradient=True,
                                 target=tgt))
            continue
        wl = filaments_workload(
            f"req-{i}-{side}x{side}-{nf}f", (side, side), nf, 0, wseed,
            node_counts=counts, amps=amps, periods=periods, steps=steps,
            overrides=_lbm_overrides(0.8, 4))
        out.append(ServeItem(arrival=t, workload=wl, steps=steps, gradient=False))
    return out


# The folding experiment's default physics (experiments/conf/experiment/
# filament_folding.yaml:53-90, assembled by experiment.py:925-941): stable
# Cosserat with the steric term at dt 0.02, 150 simulated steps.  `dt` of the
# Cosserat scheme is its inertia time step and must equal the step's dt.
FOLDING_DT = 0.02
FOLDING_STEPS = 150
FOLDING_COSSERAT = dict(cycles_per_step=10, stiffness_stretch=55000.0, stiffness_shear=1.0,
                        damping=0.98, enable_steric=True, steric_strength=1300.0,
                        steric_sigma=2.0, steric_scatter_value=1.0, steric_neighbor_skip=3)


def folding(seed=3, grid=(256, 256), n_fil=8, n_nodes=48, steps=500, modes=4, amplitude=5.0):
    """Differentiable folding: ``n_fil`` straight filaments of ``n_nodes``
    nodes (uniform start points and headings) on a ``grid`` node domain, the
    rod-only step with the reference's PBD_SCHEME_NO_FLUID parameters and the
    steric term off, and a seeded smooth target curve per filament: the
    straight rest shape displaced along its normal by a sum of ``modes``
    Fourier modes sin(pi k s + phi_k), k = 1..modes, s in [0, 1], whose
    absolute coefficients sum to ``amplitude`` lattice units.
    Returns (workload, target (N, 2) float32)."""
    wl = filaments_workload(f"folding-{grid[0]}x{grid[1]}-{n_fil}x{n_nodes}", grid, n_fil,
                            n_nodes, seed, jitter=0.0, amp=0.0, steps=steps,
                            overrides=dict(enable_steric=False))
    rng = np.random.default_rng(seed + 1000)
    h, w = grid
    tgt = []
    for f in range(n_fil):
        p = wl.position[f * n_nodes:(f + 1) * n_nodes].astype(np.float64)
        d = _min_image(grid, p[-1:], p[:1])[0].astype(np.float64)
        t = d / max(np.linalg.norm(d), 1e-12)
        nrm = np.array([-t[1], t[0]])
        coef = rng.normal(size=modes)
        coef *= amplitude / np.abs(coef).sum()
        phi = rng.uniform(0, 2 * np.pi, size=modes)
        s = np.arange(n_nodes) / max(n_nodes - 1, 1)
        off = sum(coef[k] * np.sin(np.pi * (k + 1) * s + phi[k]) for k in range(modes))
        q = p + off[:, None] * nrm[None, :]
        tgt.append(np.mod(q, np.array([w, h], np.float64)))
    return wl, np.concatenate(tgt).astype(np.float32)


def device_state(wl: Workload, device="cuda"):
    """The workload's initial state as (Nodes, Edges, Fields) of torch tensors
    on ``device`` (float32, int32 topology)."""
    import torch
    from .structs import Edges, Fields, Nodes

    def t(a, dt=torch.float32):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)

    nodes = Nodes(position=t(wl.position), velocity=t(wl.velocity),
                  color_bend=None if wl.color_bend is None else t(wl.color_bend, torch.int32),
                  debug_vector=t(np.zeros_like(wl.position)))
    edges = Edges(pairs=t(wl.pairs, torch.int32), theta=t(wl.theta),
                  rest_lengths=t(wl.rest_lengths), bending_pairs=t(wl.bending_pairs, torch.int32),
                  bending_rest_angles=t(wl.bending_rest_angles),
                  bending_stiffness=t(wl.bending_stiffness))
    fields = Fields(grid_shape=tuple(wl.grid_shape), steric=t(wl.steric),
                    fluid_velocity=t(wl.fluid_velocity), f_grid=t(wl.f_grid))
    return nodes, edges, fields
