"""The filament-folding experiment's objective over several frames
(experiments/filament_folding/experiment.py:787-916), differentiated through
the rod adjoint.

The reference scores a rollout of ``num_steps + stability_steps`` steps:
  * the shape loss (DCD or Chamfer against a target point cloud, PCA-aligned
    or centred) averaged over ``num_eval_frames`` frames spread over the
    second half of the first ``num_steps`` steps (``eval_indices``);
  * the stability term: over ``num_stability_frames`` frames of the
    stability window (``stability_indices``), the mean squared minimum-image
    displacement from the last evaluation frame and from an anchor frame
    ``stability_anchor_offset`` steps later;
  * the smooth maximum of the nearest-neighbour distances over the
    evaluation frames, and an optional smoothness penalty on consecutive
    rest angles;
with the rest angles parameterised as tanh(raw) * curvature_scale
(experiment.py:739-745).  Frame indices are the reference's timeseries
indices: index t holds the state after t + 1 steps.

The losses run in HIP kernels (``losses``); the frames are checkpoint
boundaries of the adjoint's forward rollout and their cotangents enter the
reverse pass there (``RodAdjoint.backward_frames``, ``backward_params``).
The gradient is taken with respect to the raw rest angles and, when they are
given, the raw rest lengths (distance_min + sigmoid(raw) (distance_max -
distance_min), experiment.py:741-742); otherwise the edges' rest lengths are
held fixed.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import torch

from . import losses as L
from ._runtime import as_f32
from .grad import FoldingOptimizer, RodAdjoint
from .structs import Edges, Fields, Nodes


@dataclass(frozen=True)
class FoldingLossConfig:
    """Loss settings of experiments/conf/experiment/filament_folding.yaml:36-48
    (loss section) and :66 (optimization.stability_weight).  The yaml turns
    the PCA alignment off (centring only); use_pca_rotation=True selects the
    PCA-aligned form of the losses."""
    loss_mode: str = "dcd"              # "dcd" or "pointcloud" (Chamfer)
    use_pca_rotation: bool = False
    num_eval_frames: int = 4
    num_stability_frames: int = 4
    stability_anchor_offset: int = 30
    stability_weight: float = 1.0
    max_dist_weight: float = 1.0
    smooth_max_alpha: float = 0.1
    smoothness_weight: float = 0.0
    curvature_scale: float = 0.6
    distance_min: float = 0.1
    distance_max: float = 4.0


def eval_indices(num_steps: int, num_eval_frames: int) -> list:
    """experiment.py:787-792: ``num_eval_frames`` timeseries indices evenly
    spaced from num_steps // 2 - 1 to num_steps - 1 (Python rounding)."""
    n = int(num_steps)
    k = int(num_eval_frames)
    if k == 1:
        return [n - 1]
    lo = n // 2 - 1
    return [int(round(lo + (n - 1 - lo) * i / (k - 1))) for i in range(k)]


def stability_indices(num_steps: int, stability_steps: int, num_stability_frames: int) -> list:
    """experiment.py:794-801: indices over the stability window num_steps ..
    num_steps + stability_steps - 1 (its last index alone for one frame,
    none without a window)."""
    n, w, k = int(num_steps), int(stability_steps), int(num_stability_frames)
    if w <= 0 or k <= 1:
        return [n + w - 1] if w > 0 else []
    end = n + w - 1
    return [int(round(n + (end - n) * i / (k - 1))) for i in range(k)]


def anchor_index(num_steps: int, stability_steps: int, eval_last: int, offset: int):
    """experiment.py:872-877: the anchor frame, ``offset`` indices after the
    last evaluation frame within the rollout; None when it would not lie
    after it (the reference's use_anchor false)."""
    a = min(int(eval_last) + int(offset), int(num_steps) + int(stability_steps) - 1)
    return a if a > int(eval_last) else None


def rest_angles(raw: torch.Tensor, cfg: FoldingLossConfig) -> torch.Tensor:
    """experiment.py:740: tanh(raw) * curvature_scale."""
    return torch.tanh(raw) * cfg.curvature_scale


def rest_lengths(raw: torch.Tensor, cfg: FoldingLossConfig) -> torch.Tensor:
    """experiment.py:741-742: distance_min + sigmoid(raw) (distance_max -
    distance_min)."""
    return cfg.distance_min + torch.sigmoid(raw) * (cfg.distance_max - cfg.distance_min)


def _frame_loss(fn, p, t, use_pca, **kw):
    if use_pca:
        return L.pca_aligned(fn, p, t, **kw)
    # centred branch (experiment.py:826-827): the gradient goes through the
    # particles' centring
    l, g = fn(p - p.mean(dim=0), t - t.mean(dim=0), **kw)
    return l, g - g.mean(dim=0)


def frame_objective(positions: dict, target: torch.Tensor, grid_shape, num_steps: int,
                    stability_steps: int, cfg: FoldingLossConfig):
    """The position-dependent part of experiment.py:803-916 (shape +
    stability_weight * stability + max_dist_weight * smooth max) from
    ``positions``, which maps a step count (timeseries index + 1) to the
    positions after that many steps and must hold every frame
    ``objective_steps`` names.  Returns (total, shape, stability, {step:
    d total / d positions}); the scalars are device tensors."""
    if cfg.loss_mode not in ("dcd", "pointcloud"):
        raise ValueError(f"loss_mode {cfg.loss_mode!r}: only 'dcd' and 'pointcloud' are built")
    ev = [t + 1 for t in eval_indices(num_steps, cfg.num_eval_frames)]
    st = [t + 1 for t in stability_indices(num_steps, stability_steps, cfg.num_stability_frames)]
    an = anchor_index(num_steps, stability_steps, ev[-1] - 1, cfg.stability_anchor_offset)
    an = None if an is None else an + 1
    fn = L.chamfer_loss if cfg.loss_mode == "pointcloud" else L.dcd_loss
    some = positions[ev[0]]
    bars = {t: torch.zeros_like(positions[t]) for t in set(ev) | set(st) | ({an} - {None})}
    ne = len(ev)
    shape = torch.zeros((), dtype=torch.float32, device=some.device)
    for t in ev:
        l, g = _frame_loss(fn, positions[t], target, cfg.use_pca_rotation)
        shape = shape + l / ne
        bars[t] += g / ne
    main = ev[-1]
    if st:
        fr = torch.stack([positions[t] for t in st])
        stab, gf, gm, ga = L.frame_stability(grid_shape, fr, positions[main],
                                             None if an is None else positions[an])
        w = cfg.stability_weight
        for i, t in enumerate(st):
            bars[t] += w * gf[i]
        bars[main] += w * gm
        if an is not None:
            bars[an] += w * ga
    else:
        # experiment.py:891-892: the last frame against itself
        stab = torch.zeros((), dtype=torch.float32, device=some.device)
    total = shape + cfg.stability_weight * stab
    if cfg.max_dist_weight > 0.0:
        md = torch.zeros((), dtype=torch.float32, device=some.device)
        for t in ev:
            l, g = _frame_loss(L.smooth_max_nn, positions[t], target, cfg.use_pca_rotation,
                               alpha=cfg.smooth_max_alpha)
            md = md + l / ne
            bars[t] += cfg.max_dist_weight * g / ne
        total = total + cfg.max_dist_weight * md
    return total, shape, stab, bars


def objective_steps(num_steps: int, stability_steps: int, cfg: FoldingLossConfig) -> list:
    """Step counts of every frame the objective reads (evaluation,
    stability and anchor frames)."""
    ev = eval_indices(num_steps, cfg.num_eval_frames)
    st = stability_indices(num_steps, stability_steps, cfg.num_stability_frames)
    an = anchor_index(num_steps, stability_steps, ev[-1], cfg.stability_anchor_offset)
    return sorted({t + 1 for t in ev + st + ([] if an is None else [an])})


def folding_value_and_grad(nodes: Nodes, edges: Edges, fields: Fields, solver_config, dt: float,
                           num_steps: int, stability_steps: int, raw_angles: torch.Tensor,
                           target: torch.Tensor, cfg: FoldingLossConfig = FoldingLossConfig(),
                           checkpoint_every: int = 50, raw_lengths: torch.Tensor = None):
    """One evaluation of the folding experiment's loss (experiment.py:803-916
    for one initial state) and its gradient with respect to the raw rest
    angles: rolls out num_steps + stability_steps steps from (nodes, edges,
    fields) with rest angles tanh(raw) * curvature_scale, scores the frames
    and runs the adjoint back through them.  Returns (total, shape,
    stability, d total / d raw_angles); with ``raw_lengths`` (E,), the rest
    lengths are ``rest_lengths(raw_lengths)`` and the tuple ends with
    d total / d raw_lengths as well."""
    dev = nodes.position.device
    raw = as_f32(raw_angles, dev)
    rest = rest_angles(raw, cfg)
    e = replace(edges, bending_rest_angles=rest)
    if raw_lengths is not None:
        raw_l = as_f32(raw_lengths, dev)
        e = replace(e, rest_lengths=rest_lengths(raw_l, cfg))
    ad = RodAdjoint(nodes, e, fields, solver_config, dt)
    total_steps = int(num_steps) + int(stability_steps)
    steps = objective_steps(num_steps, stability_steps, cfg)
    ad.forward(total_steps, checkpoint_every, frames=steps)
    pos = {t: ad.frame_position(t) for t in steps}
    tgt = as_f32(target, dev)
    total, shape, stab, bars = frame_objective(pos, tgt, ad.grid_shape, num_steps,
                                               stability_steps, cfg)
    if raw_lengths is None:
        g_rest = ad.backward_frames(bars)[0]
    else:
        g_rest, g_len = ad.backward_params(bars)[:2]
    if cfg.smoothness_weight > 0.0 and rest.shape[0] > 1:
        # experiment.py:895-896: mean((a[1:] - a[:-1])^2)
        d = rest[1:] - rest[:-1]
        total = total + cfg.smoothness_weight * (d * d).mean()
        gd = cfg.smoothness_weight * 2.0 * d / d.shape[0]
        g_rest = g_rest.clone()
        g_rest[1:] += gd
        g_rest[:-1] -= gd
    th = torch.tanh(raw)
    g_raw = g_rest * cfg.curvature_scale * (1.0 - th * th)
    if raw_lengths is None:
        return total, shape, stab, g_raw
    sg = torch.sigmoid(raw_l)
    return total, shape, stab, g_raw, g_len * (cfg.distance_max - cfg.distance_min) * sg * (1.0 - sg)


def batch_folding_value_and_grad(instances, solver_config, dt: float, num_steps: int,
                                 stability_steps: int, raw_angles: torch.Tensor,
                                 target: torch.Tensor,
                                 cfg: FoldingLossConfig = FoldingLossConfig(),
                                 checkpoint_every: int = 50, raw_lengths: torch.Tensor = None):
    """The experiment's batched loss (experiment.py:918-923: the mean over
    ``optimization.batch_size`` initial states of the per-state loss, with
    shared parameters): one batched rollout and adjoint
    (``BatchRodAdjoint``) for every instance of ``instances`` ((Nodes, Edges,
    Fields) of the same filament topology).  Returns (mean total, mean
    shape, mean stability, d mean / d raw_angles[, d mean / d raw_lengths])
    as ``folding_value_and_grad``."""
    from .grad import BatchRodAdjoint
    dev = instances[0][0].position.device
    raw = as_f32(raw_angles, dev)
    rest = rest_angles(raw, cfg)
    raw_l = None if raw_lengths is None else as_f32(raw_lengths, dev)
    inst = []
    for nodes, edges, fields in instances:
        e = replace(edges, bending_rest_angles=rest)
        if raw_l is not None:
            e = replace(e, rest_lengths=rest_lengths(raw_l, cfg))
        inst.append((nodes, e, fields))
    ad = BatchRodAdjoint(inst, solver_config, dt)
    total_steps = int(num_steps) + int(stability_steps)
    steps = objective_steps(num_steps, stability_steps, cfg)
    ad.forward(total_steps, checkpoint_every, frames=steps)
    pos = {t: ad.frame_positions(t) for t in steps}
    tgt = as_f32(target, dev)
    nb = len(instances)
    outs = []
    for i, (_, _, fields) in enumerate(inst):
        outs.append(frame_objective({t: pos[t][i] for t in steps}, tgt,
                                    tuple(fields.grid_shape), num_steps, stability_steps, cfg))
    # the mean over the batch: each instance's cotangents scaled by 1 / nb
    bars = {t: [o[3][t] / nb for o in outs] for t in steps}
    grads = ad.backward_params(bars)
    total = sum(o[0] for o in outs) / nb
    shape = sum(o[1] for o in outs) / nb
    stab = sum(o[2] for o in outs) / nb
    g_rest = sum(g[0] for g in grads)
    if cfg.smoothness_weight > 0.0 and rest.shape[0] > 1:
        d = rest[1:] - rest[:-1]
        total = total + cfg.smoothness_weight * (d * d).mean()
        gd = cfg.smoothness_weight * 2.0 * d / d.shape[0]
        g_rest = g_rest.clone()
        g_rest[1:] += gd
        g_rest[:-1] -= gd
    th = torch.tanh(raw)
    g_raw = g_rest * cfg.curvature_scale * (1.0 - th * th)
    if raw_l is None:
        return total, shape, stab, g_raw
    sg = torch.sigmoid(raw_l)
    g_len = sum(g[1] for g in grads)
    return total, shape, stab, g_raw, g_len * (cfg.distance_max - cfg.distance_min) * sg * (1.0 - sg)


def optimize_folding_objective(nodes: Nodes, edges: Edges, fields: Fields, solver_config,
                               dt: float, num_steps: int, stability_steps: int,
                               target: torch.Tensor, iterations: int,
                               cfg: FoldingLossConfig = FoldingLossConfig(), raw_angles=None,
                               optimizer: FoldingOptimizer = None, checkpoint_every: int = 50,
                               raw_lengths=None):
    """The folding experiment's loop over its parameters
    (experiment.py:397-403, 985-1030 for one initial state, without
    curriculum stages, noise or visualisation): the raw rest angles and,
    when ``raw_lengths`` is given (the reference starts both at zero), the
    raw rest lengths, updated together by ``FoldingOptimizer`` (the global
    norm is taken over both, as optax's over the parameter tree).  Returns
    (raw angles, raw lengths or None, per-iteration total losses)."""
    dev = nodes.position.device
    nb = edges.bending_rest_angles.shape[0]
    raw = (torch.zeros(nb, dtype=torch.float32, device=dev)
           if raw_angles is None else as_f32(raw_angles, dev).clone())
    if raw_lengths is not None:
        raw = torch.cat([raw, as_f32(raw_lengths, dev)])
    opt = optimizer or FoldingOptimizer(total_steps=iterations)
    hist = []
    for _ in range(iterations):
        out = folding_value_and_grad(nodes, edges, fields, solver_config, dt, num_steps,
                                     stability_steps, raw[:nb], target, cfg, checkpoint_every,
                                     None if raw_lengths is None else raw[nb:])
        hist.append(float(out[0].item()))
        g = out[3] if raw_lengths is None else torch.cat([out[3], out[4]])
        raw = opt.update(raw, g)
    return raw[:nb], (None if raw_lengths is None else raw[nb:]), hist
