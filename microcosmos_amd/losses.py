"""Folding losses of the filament-folding experiment on the device
(experiments/filament_folding/experiment.py:57-165), each returning the loss
and its gradient with respect to the particle positions -- the cotangent the
rod adjoint (``grad.RodAdjoint.backward``) takes.

The nearest-neighbour search, the density counts, the loss and the gradient
run in HIP kernels (``mc_point_loss``); the PCA alignment's 2 x 2 principal
axes and the rotation of the centred cloud are a handful of elementwise device
operations around them.
"""
from __future__ import annotations

import math

import torch

from . import _lib
from ._lib import check
from ._runtime import as_f32, ptr, stream_handle

CHAMFER, DCD, SMOOTH_MAX = 0, 1, 2


def _point_loss(kind: int, particles: torch.Tensor, targets: torch.Tensor, alpha: float = 0.0):
    d = particles.device
    if d.type != "cuda":
        raise RuntimeError("microcosmos_amd runs on the GPU only")
    lib = _lib.load()
    p = as_f32(particles, d)
    t = as_f32(targets, d)
    n, m = p.shape[0], t.shape[0]
    ws = torch.empty(int(lib.mc_point_loss_workspace(n, m)), dtype=torch.uint8, device=d)
    loss = torch.empty((), dtype=torch.float32, device=d)
    grad = torch.empty((n, 2), dtype=torch.float32, device=d)
    check(lib.mc_point_loss(kind, ptr(p), n, ptr(t), m, float(alpha), ptr(loss), ptr(grad),
                            ptr(ws), stream_handle(d)))
    return loss, grad


def chamfer_loss(particles: torch.Tensor, targets: torch.Tensor):
    """experiment.py:57-63 -- mean_i min_j |p_i - t_j|^2 + mean_j min_i
    |p_i - t_j|^2; returns (loss, d loss / d particles)."""
    return _point_loss(CHAMFER, particles, targets)


def dcd_loss(particles: torch.Tensor, targets: torch.Tensor):
    """experiment.py:105-128 -- density-aware Chamfer: each nearest-neighbour
    term divided by how many points chose the same neighbour."""
    return _point_loss(DCD, particles, targets)


def smooth_max_nn(particles: torch.Tensor, targets: torch.Tensor, alpha: float = 0.1):
    """experiment.py:139-149 -- logsumexp(alpha sqrt(min_j d2 + 1e-8)) / alpha."""
    return _point_loss(SMOOTH_MAX, particles, targets, alpha)


def _principal_angle(x: torch.Tensor) -> torch.Tensor:
    # direction of the eigenvector of the larger eigenvalue of the 2 x 2
    # covariance + 1e-6 I (experiment.py:74-81), up to its sign: half the
    # angle of (2 c_xy, c_xx - c_yy)
    c = x - x.mean(dim=0)
    cxx = (c[:, 0] * c[:, 0]).mean() + 1e-6
    cyy = (c[:, 1] * c[:, 1]).mean() + 1e-6
    cxy = (c[:, 0] * c[:, 1]).mean()
    return 0.5 * torch.atan2(2.0 * cxy, cxx - cyy)


def pca_rotation_angle(particles: torch.Tensor, targets: torch.Tensor) -> float:
    """experiment.py:66-84: the targets' principal-axis angle minus the
    particles' (each defined up to pi, which the callers' +-pi minimum
    absorbs)."""
    a = _principal_angle(as_f32(targets, targets.device).double()) - \
        _principal_angle(as_f32(particles, particles.device).double())
    return float(a.item())


def pca_aligned(loss_fn, particles: torch.Tensor, targets: torch.Tensor, **kw):
    """The *_pca_aligned wrappers (experiment.py:95-102, 131-136, 152-160):
    the loss of the centred particles rotated by the principal-axis angle and
    by the angle + pi against the centred targets, the smaller one kept (a
    tie averages the two gradients, as jnp.minimum); the angle carries no
    gradient, the rotation and the centring do.  Stays on the device: the
    angle, its cosine and sine and the choice are tensors, so no step waits
    for the host."""
    p = as_f32(particles, particles.device)
    t = as_f32(targets, targets.device)
    a = _principal_angle(t.double()) - _principal_angle(p.double())
    tc = t - t.mean(dim=0)
    c = p - p.mean(dim=0)
    out = []
    for k in (0, 1):
        ang = a + k * math.pi
        co, si = torch.cos(ang).float(), torch.sin(ang).float()
        q = torch.stack([c[:, 0] * co - c[:, 1] * si, c[:, 0] * si + c[:, 1] * co], dim=-1)
        l, gq = loss_fn(q, tc, **kw)
        gc = torch.stack([gq[:, 0] * co + gq[:, 1] * si, -gq[:, 0] * si + gq[:, 1] * co], dim=-1)
        out.append((l, gc - gc.mean(dim=0)))
    (l0, g0), (l1, g1) = out
    g = torch.where(l1 < l0, g1, torch.where(l0 < l1, g0, 0.5 * (g0 + g1)))
    return torch.minimum(l0, l1), g


def frame_stability(grid_shape, frames: torch.Tensor, main: torch.Tensor,
                    anchor: torch.Tensor = None):
    """experiment.py:878-893: the stability term, mean over the frames
    (S, n, 2) of mean(d(x_s, main)^2) + mean(d(x_s, anchor)^2), d the
    minimum-image displacement on the (H, W) torus (mc_frame_stability).
    ``anchor`` None leaves its term out.  Returns (loss, d/d frames,
    d/d main, d/d anchor or None)."""
    d = frames.device
    if d.type != "cuda":
        raise RuntimeError("microcosmos_amd runs on the GPU only")
    lib = _lib.load()
    fr = as_f32(frames, d)
    s, n = fr.shape[0], fr.shape[1]
    m = as_f32(main, d)
    a = as_f32(anchor, d) if anchor is not None else None
    if fr.shape != (s, n, 2) or m.shape != (n, 2) or (a is not None and a.shape != (n, 2)):
        raise ValueError("frames must be (S, n, 2) and main / anchor (n, 2)")
    h, w = grid_shape
    loss = torch.empty((), dtype=torch.float32, device=d)
    gf = torch.empty_like(fr)
    gm = torch.empty_like(m)
    ga = torch.empty_like(a) if a is not None else None
    check(lib.mc_frame_stability(ptr(fr), s, n, ptr(m), ptr(a), int(h), int(w), ptr(loss),
                                 ptr(gf), ptr(gm), ptr(ga), stream_handle(d)))
    return loss, gf, gm, ga
