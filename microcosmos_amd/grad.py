"""Gradients of rod rollouts through hand-written adjoint kernels
(mc_rod_adjoint_forward / mc_rod_adjoint_backward).

The reference differentiates its rollouts with ``jax.value_and_grad``
through ``lax.scan`` of ``simulate.step``
(experiments/filament_folding/experiment.py:985-1018), storing every step.
Here the forward rollout keeps a checkpoint every ``checkpoint_every`` steps
and the reverse pass recomputes each segment before running the per-step
adjoint kernel backwards through it, so the device memory of a gradient is
(checkpoints + one segment) x state instead of the whole trajectory.

Scope: the step without the fluid (PBD or stable Cosserat constraint,
``enable_fluid=False``, with or without the steric term), single or batched,
and the fluid-coupled step (``enable_fluid=True``, single-instance plans:
the immersed boundary with the node positions held constant, as the
reference's stop_gradient does at fluid.py:239, and the TRT/Guo lattice),
with the bending rest angles (and, through ``backward_params``, the rest
lengths) as the parameters.  The steric term and the fluid couple every
filament each step, so with either every step's state is kept
(``checkpoint_every`` is 1; with the fluid the step's input lattice too) and
each reverse step recomputes that step's rod half.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import replace

import numpy as np
import torch

from . import _lib
from ._lib import McState, check
from ._runtime import Plan, as_f32, edge_params, params_from_config, ptr, stream_handle
from .structs import Edges, Fields, Nodes


def hinge_center_nodes(pairs, bending_pairs) -> np.ndarray:
    """Node at the centre of every hinge: the target node of its incoming edge
    (examples/01_es_locomotion.ipynb ``bending_center_nodes``)."""
    pa = np.asarray(pairs).reshape(-1, 2)
    bp = np.asarray(bending_pairs).reshape(-1, 2)
    return pa[bp[:, 0], 1].astype(np.int64)


def node_gradient(rest_angle_grad: torch.Tensor, centers, num_nodes: int) -> torch.Tensor:
    """Gradient with respect to per-node rest curvatures kappa_n when hinge b
    uses kappa at its centre node (rest_b = kappa[center_b])."""
    idx = torch.as_tensor(np.asarray(centers), device=rest_angle_grad.device)
    out = torch.zeros(num_nodes, dtype=rest_angle_grad.dtype, device=rest_angle_grad.device)
    return out.index_add_(0, idx, rest_angle_grad)


def min_image(grid_shape, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """utils.py:18-23 minimum-image displacement a - b on the (H, W) torus."""
    h, w = grid_shape
    s = torch.tensor([float(w), float(h)], dtype=a.dtype, device=a.device)
    return torch.remainder(a - b + s / 2, s) - s / 2


class RodAdjoint:
    """A rollout of the step that can be differentiated: the rod-only step, or
    with ``enable_fluid`` the fluid-coupled step (immersed boundary with the
    node positions held constant, as the reference's stop_gradient does, and
    the TRT/Guo lattice; single-instance plans, a checkpoint every step).

    ``forward(num_steps)`` advances a private copy of the state and records
    checkpoints; ``backward(position_bar)`` returns d loss / d rest angles
    (and the gradient with respect to the initial state) for a loss of the
    final positions (and velocities).
    """

    def __init__(self, nodes: Nodes, edges: Edges, fields: Fields, solver_config, dt: float):
        if getattr(solver_config, "solver_type", "pbd") not in ("pbd", "cosserat"):
            raise NotImplementedError("rod adjoint: PBD or stable Cosserat constraint")
        d = nodes.position.device
        if d.type != "cuda":
            raise RuntimeError("microcosmos_amd runs on the GPU only")
        self.dev = d
        self.params = params_from_config(solver_config)
        self.dt = float(dt)
        self.position = as_f32(nodes.position, d).clone()
        self.velocity = as_f32(nodes.velocity, d).clone()
        self.debug_vector = as_f32(nodes.debug_vector, d).clone()
        self.theta = as_f32(edges.theta, d).clone()
        self.grid_shape = tuple(fields.grid_shape)
        self.steric = as_f32(fields.steric, d).clone()
        # with the fluid the forward advances the lattice: private copies
        self.fluid_velocity = as_f32(fields.fluid_velocity, d).clone()
        self.f_grid = as_f32(fields.f_grid, d).clone()
        from ._runtime import _host_topology
        pa = _host_topology(edges.pairs)[0]
        bp = _host_topology(edges.bending_pairs)[0]
        # a private plan: it holds this rollout's checkpoints
        self.plan = Plan(self.grid_shape, tuple(self.f_grid.shape[1:]), self.position.shape[0],
                         pa, bp)
        self.ep, self._keep = edge_params(edges, d)
        self.num_edges = pa.shape[0]
        self.num_hinges = bp.shape[0]
        self.steps = 0

    def _state(self) -> McState:
        return McState(ptr(self.position), ptr(self.velocity), ptr(self.debug_vector),
                       ptr(self.theta), ptr(self.steric), ptr(self.fluid_velocity),
                       ptr(self.f_grid))

    def forward(self, num_steps: int, checkpoint_every: int = 50, frames=None) -> torch.Tensor:
        """Runs the rollout; ``frames`` (step counts) are the intermediate
        states a multi-frame loss will score: the checkpoint interval is
        lowered to a common divisor so each frame ends a segment."""
        lib = _lib.load()
        st = self._state()
        every = int(checkpoint_every)
        for t in frames or ():
            if not 1 <= int(t) <= int(num_steps):
                raise ValueError(f"frame {t} outside 1..{num_steps}")
            if int(t) != int(num_steps):
                every = math.gcd(every, int(t))
        check(lib.mc_rod_adjoint_forward(self.plan.handle, ctypes.byref(self.params),
                                         ctypes.byref(st), ctypes.byref(self.ep), self.dt,
                                         int(num_steps), every, stream_handle(self.dev)))
        self.steps = int(num_steps)
        return self.position

    def frame_position(self, step: int) -> torch.Tensor:
        """Positions (N, 2) after ``step`` steps of the recorded rollout (a
        frame given to ``forward``, or the last step)."""
        if int(step) == self.steps:
            return self.position
        lib = _lib.load()
        out = torch.empty_like(self.position)
        check(lib.mc_rod_adjoint_state(self.plan.handle, int(step), ptr(out), None, None,
                                       stream_handle(self.dev)))
        return out

    def backward_frames(self, frame_bars: dict, velocity_bar: torch.Tensor = None):
        """Gradient of a loss over several frames: ``frame_bars`` maps a step
        count (given to ``forward`` as a frame, or the last step) to d loss /
        d positions after that many steps.  Returns (rest_angle_grad,
        position0_grad, velocity0_grad, theta0_grad)."""
        lib = _lib.load()
        d = self.dev
        steps = np.ascontiguousarray(sorted(frame_bars), dtype=np.int64)
        bars = torch.stack([as_f32(frame_bars[int(t)], d).reshape(-1, 2) for t in steps]).contiguous()
        vb = as_f32(velocity_bar, d) if velocity_bar is not None else None
        n = self.position.shape[0]
        rg = torch.empty(self.num_hinges, dtype=torch.float32, device=d)
        g_x = torch.empty((n, 2), dtype=torch.float32, device=d)
        g_v = torch.empty((n, 2), dtype=torch.float32, device=d)
        g_t = torch.empty(self.num_edges, dtype=torch.float32, device=d)
        check(lib.mc_rod_adjoint_backward_frames(
            self.plan.handle, ctypes.byref(self.params), ctypes.byref(self.ep), self.dt,
            int(steps.shape[0]), steps.ctypes.data, ptr(bars), ptr(vb), ptr(rg), ptr(g_x),
            ptr(g_v), ptr(g_t), stream_handle(d)))
        return rg, g_x, g_v, g_t

    def backward_params(self, frame_bars: dict, velocity_bar: torch.Tensor = None):
        """``backward_frames`` that also returns the gradient with respect to
        the rest lengths (mc_rod_adjoint_backward_params): (rest_angle_grad
        (B,), rest_length_grad (E,), position0_grad, velocity0_grad,
        theta0_grad)."""
        lib = _lib.load()
        d = self.dev
        steps = np.ascontiguousarray(sorted(frame_bars), dtype=np.int64)
        bars = torch.stack([as_f32(frame_bars[int(t)], d).reshape(-1, 2) for t in steps]).contiguous()
        vb = as_f32(velocity_bar, d) if velocity_bar is not None else None
        n = self.position.shape[0]
        rg = torch.empty(self.num_hinges, dtype=torch.float32, device=d)
        lg = torch.empty(self.num_edges, dtype=torch.float32, device=d)
        g_x = torch.empty((n, 2), dtype=torch.float32, device=d)
        g_v = torch.empty((n, 2), dtype=torch.float32, device=d)
        g_t = torch.empty(self.num_edges, dtype=torch.float32, device=d)
        check(lib.mc_rod_adjoint_backward_params(
            self.plan.handle, ctypes.byref(self.params), ctypes.byref(self.ep), self.dt,
            int(steps.shape[0]), steps.ctypes.data, ptr(bars), ptr(vb), ptr(rg), ptr(lg),
            ptr(g_x), ptr(g_v), ptr(g_t), stream_handle(d)))
        return rg, lg, g_x, g_v, g_t

    def backward(self, position_bar: torch.Tensor, velocity_bar: torch.Tensor = None):
        """Returns (rest_angle_grad (B,), position0_grad (N, 2),
        velocity0_grad (N, 2), theta0_grad (E,))."""
        lib = _lib.load()
        d = self.dev
        pb = as_f32(position_bar, d)
        vb = as_f32(velocity_bar, d) if velocity_bar is not None else None
        n = self.position.shape[0]
        rg = torch.empty(self.num_hinges, dtype=torch.float32, device=d)
        g_x = torch.empty((n, 2), dtype=torch.float32, device=d)
        g_v = torch.empty((n, 2), dtype=torch.float32, device=d)
        g_t = torch.empty(self.num_edges, dtype=torch.float32, device=d)
        check(lib.mc_rod_adjoint_backward(self.plan.handle, ctypes.byref(self.params),
                                          ctypes.byref(self.ep), self.dt, ptr(pb), ptr(vb),
                                          ptr(rg), ptr(g_x), ptr(g_v), ptr(g_t),
                                          stream_handle(d)))
        return rg, g_x, g_v, g_t


def rollout_value_and_grad(nodes: Nodes, edges: Edges, fields: Fields, solver_config, dt: float,
                           num_steps: int, target: torch.Tensor, checkpoint_every: int = 50,
                           loss_fn=None):
    """Folding loss after ``num_steps`` steps and its gradient with respect
    to the (time-constant) bending rest angles.  The loss is the mean over
    nodes of the squared minimum-image distance to ``target`` (N, 2), or
    ``loss_fn(positions, target) -> (loss, d loss / d positions)`` -- e.g.
    ``losses.chamfer_loss`` or ``losses.dcd_loss`` against a target point
    cloud (M, 2)."""
    ad = RodAdjoint(nodes, edges, fields, solver_config, dt)
    x = ad.forward(num_steps, checkpoint_every)
    tgt = as_f32(target, ad.dev)
    if loss_fn is not None:
        loss, gx = loss_fn(x, tgt)
        return loss, ad.backward(gx)[0]
    dvec = min_image(ad.grid_shape, x, tgt)
    n = x.shape[0]
    loss = (dvec * dvec).sum() / n
    g = ad.backward(2.0 * dvec / n)[0]
    return loss, g


def folding_schedule(step: int, lr: float = 0.06, lr_min: float = 3e-5, warmup_steps: int = 15,
                     total_steps: int = 150) -> float:
    """Learning rate of the folding experiment without curriculum stages
    (experiments/filament_folding/experiment.py:340-395): a linear warmup
    from lr / 100 to lr over ``warmup_steps``, then a cosine decay from lr to
    lr_min over the remaining steps (optax linear_schedule, cosine_decay_schedule
    with alpha = lr_min / lr, joined at the warmup boundary).  Defaults from
    experiments/conf/experiment/filament_folding.yaml:59-66."""
    if step < warmup_steps:
        return lr * 0.01 + (lr - lr * 0.01) * (step / warmup_steps)
    decay = max(total_steps - warmup_steps, 1)
    t = min(step - warmup_steps, decay)
    alpha = lr_min / lr
    return lr * ((1.0 - alpha) * 0.5 * (1.0 + math.cos(math.pi * t / decay)) + alpha)


class FoldingOptimizer:
    """The folding experiment's update (experiment.py:397-400, 1021-1023):
    gradient clipping by global norm, then Adam (b1 0.9, b2 0.999, eps 1e-8)
    scaled by ``folding_schedule`` at the update count.  Parameters and
    gradients are device tensors; the update is elementwise."""

    def __init__(self, lr: float = 0.06, lr_min: float = 3e-5, warmup_steps: int = 15,
                 total_steps: int = 150, grad_clip: float = 1.0, b1: float = 0.9,
                 b2: float = 0.999, eps: float = 1e-8):
        self.lr, self.lr_min, self.warmup, self.total = lr, lr_min, warmup_steps, total_steps
        self.clip, self.b1, self.b2, self.eps = grad_clip, b1, b2, eps
        self.count = 0
        self.mu = None
        self.nu = None

    def update(self, params: torch.Tensor, grad: torch.Tensor) -> torch.Tensor:
        g = grad.to(torch.float32)
        norm = torch.sqrt((g * g).sum())
        g = torch.where(norm < self.clip, g, g / norm * self.clip)
        if self.mu is None:
            self.mu = torch.zeros_like(g)
            self.nu = torch.zeros_like(g)
        self.mu = self.b1 * self.mu + (1.0 - self.b1) * g
        self.nu = self.b2 * self.nu + (1.0 - self.b2) * (g * g)
        t = self.count + 1
        mu_hat = self.mu / (1.0 - self.b1 ** t)
        nu_hat = self.nu / (1.0 - self.b2 ** t)
        lr = folding_schedule(self.count, self.lr, self.lr_min, self.warmup, self.total)
        self.count += 1
        return params - lr * mu_hat / (torch.sqrt(nu_hat) + self.eps)


def optimize_folding(nodes: Nodes, edges: Edges, fields: Fields, solver_config, dt: float,
                     num_steps: int, target: torch.Tensor, iterations: int,
                     optimizer: FoldingOptimizer = None, checkpoint_every: int = 50,
                     loss_fn=None):
    """The folding experiment's optimisation loop over the bending rest angles
    (experiment.py:985-1030 without curriculum, noise or visualisation): each
    iteration rolls the rod out from the same initial state, takes the loss of
    the final positions and its gradient through the adjoint kernels, and
    applies ``optimizer``; ``loss_fn`` as in ``rollout_value_and_grad``.
    Returns (rest angles, per-iteration losses)."""
    opt = optimizer or FoldingOptimizer(total_steps=iterations)
    rest = as_f32(edges.bending_rest_angles, nodes.position.device).clone()
    losses = []
    for _ in range(iterations):
        e = replace(edges, bending_rest_angles=rest)
        loss, g = rollout_value_and_grad(nodes, e, fields, solver_config, dt, num_steps, target,
                                         checkpoint_every, loss_fn)
        losses.append(float(loss.item()))
        rest = opt.update(rest, g)
    return rest, losses


class BatchRodAdjoint:
    """Rod adjoints of several independent instances in one batched plan
    (the batch the folding experiment vmaps over, experiment.py:915-923): the
    same forward and reverse launches carry every instance's filaments, so a
    batch costs about as much time as one instance while the kernels stay
    latency-bound.  ``instances`` is a list of (Nodes, Edges, Fields); every
    instance's gradient equals the one it gets alone (the kernels never mix
    instances)."""

    def __init__(self, instances, solver_config, dt: float):
        from .batch import Batch
        if solver_config.enable_fluid or \
                getattr(solver_config, "solver_type", "pbd") not in ("pbd", "cosserat"):
            raise NotImplementedError("rod adjoint: PBD or stable Cosserat constraint with "
                                      "enable_fluid=False")
        self.batch = Batch(instances, solver_config, dt)
        if self.batch.dev.type != "cuda":
            raise RuntimeError("microcosmos_amd runs on the GPU only")
        self.dt = float(dt)
        self.steps = 0

    def forward(self, num_steps: int, checkpoint_every: int = 50, frames=None):
        """Advances every instance ``num_steps`` steps; returns the list of
        final (N_i, 2) positions.  ``frames`` as in ``RodAdjoint.forward``."""
        b = self.batch
        lib = _lib.load()
        st = b._state()
        for t in frames or ():
            if not 1 <= int(t) <= int(num_steps):
                raise ValueError(f"frame {t} outside 1..{num_steps}")
            if int(t) != int(num_steps):
                checkpoint_every = math.gcd(int(checkpoint_every), int(t))
        ep = _lib.McEdgeParams(ptr(b.rest_lengths), ptr(b.rest_angles), ptr(b.stiffness))
        self._ep = ep
        check(lib.mc_rod_adjoint_forward(b.plan.handle, ctypes.byref(b.params), ctypes.byref(st),
                                         ctypes.byref(ep), self.dt, int(num_steps),
                                         int(checkpoint_every), stream_handle(b.dev)))
        self.steps = int(num_steps)
        x = b.position.view(-1, 2)
        return [x[int(b.node_off[i]):int(b.node_off[i + 1])] for i in range(b.n)]

    def backward(self, position_bars):
        """position_bars: per instance d loss_i / d final positions (N_i, 2).
        Returns the per-instance rest-angle gradients (B_i,)."""
        b = self.batch
        lib = _lib.load()
        pb = torch.cat([as_f32(g, b.dev).reshape(-1) for g in position_bars]).contiguous()
        rg = torch.empty(int(b.hinge_off[-1]), dtype=torch.float32, device=b.dev)
        check(lib.mc_rod_adjoint_backward(b.plan.handle, ctypes.byref(b.params),
                                          ctypes.byref(self._ep), self.dt, ptr(pb), None, ptr(rg),
                                          None, None, None, stream_handle(b.dev)))
        return [rg[int(b.hinge_off[i]):int(b.hinge_off[i + 1])] for i in range(b.n)]

    def frame_positions(self, step: int):
        """Per instance, the positions (N_i, 2) after ``step`` steps of the
        recorded rollout (a frame given to ``forward``, or the last step)."""
        b = self.batch
        if int(step) == self.steps:
            x = b.position.view(-1, 2)
        else:
            x = torch.empty((int(b.node_off[-1]), 2), dtype=torch.float32, device=b.dev)
            check(_lib.load().mc_rod_adjoint_state(b.plan.handle, int(step), ptr(x), None, None,
                                                   stream_handle(b.dev)))
        return [x[int(b.node_off[i]):int(b.node_off[i + 1])] for i in range(b.n)]

    def backward_params(self, frame_bars: dict):
        """``frame_bars`` maps a frame step to the per-instance cotangents
        (list of (N_i, 2)).  Returns per instance (rest-angle gradient (B_i,),
        rest-length gradient (E_i,))."""
        b = self.batch
        steps = np.ascontiguousarray(sorted(frame_bars), dtype=np.int64)
        bars = torch.stack([torch.cat([as_f32(g, b.dev).reshape(-1) for g in frame_bars[int(t)]])
                            for t in steps]).contiguous()
        rg = torch.empty(int(b.hinge_off[-1]), dtype=torch.float32, device=b.dev)
        lg = torch.empty(int(b.edge_off[-1]), dtype=torch.float32, device=b.dev)
        check(_lib.load().mc_rod_adjoint_backward_params(
            b.plan.handle, ctypes.byref(b.params), ctypes.byref(self._ep), self.dt,
            int(steps.shape[0]), steps.ctypes.data, ptr(bars), None, ptr(rg), ptr(lg), None, None,
            None, stream_handle(b.dev)))
        return [(rg[int(b.hinge_off[i]):int(b.hinge_off[i + 1])],
                 lg[int(b.edge_off[i]):int(b.edge_off[i + 1])]) for i in range(b.n)]


def batch_value_and_grad(instances, solver_config, dt: float, num_steps: int, targets,
                         checkpoint_every: int = 50, loss_fn=None):
    """Per-instance folding losses and rest-angle gradients of a batch of
    independent rollouts (one batched adjoint).  ``targets`` holds one target
    per instance; ``loss_fn`` as in ``rollout_value_and_grad`` (default: the
    mean squared minimum-image distance).  The folding experiment's batched
    objective is the mean of the losses, whose gradient with respect to shared
    rest angles is the mean of these gradients."""
    ad = BatchRodAdjoint(instances, solver_config, dt)
    xs = ad.forward(num_steps, checkpoint_every)
    losses, bars = [], []
    for (nodes, _, fields), x, t in zip(instances, xs, targets):
        tgt = as_f32(t, x.device)
        if loss_fn is not None:
            loss, gx = loss_fn(x, tgt)
        else:
            dvec = min_image(tuple(fields.grid_shape), x, tgt)
            n = x.shape[0]
            loss, gx = (dvec * dvec).sum() / n, 2.0 * dvec / n
        losses.append(loss)
        bars.append(gx)
    return losses, ad.backward(bars)
