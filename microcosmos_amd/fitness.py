"""Swimmer fitness: centre of mass and displacement over a rollout.

Mirrors the reference callers of the step:
  * ``periodic_com`` / ``DisplacementTracker`` --
    experiments/swim_sweep/swimmers/swim_base.py:167-209 (reference-node
    minimum-image centre, cumulative per-step minimum-image displacement,
    optional ``displacement_start_step``, stability check);
  * ``rollout_fitness`` -- examples/01_es_locomotion.ipynb ``rollout_fitness``
    (per-creature plain mean, x displacement summed over creatures);
  * ``com_bins`` -- utils.py:25-27 ``index_from_position`` of the centres.

Every centre is computed on the device (``mc_center_of_mass``); inside a
``Rollout`` the per-step update runs in the step's own stream (and graph), so a
whole rollout's fitness costs no host round trip.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from ._runtime import as_f32, ptr, stream_handle


def _offsets(groups, n, device):
    """Group offsets (G + 1,) int32 from a list of slices / lengths / offsets."""
    if groups is None:
        off = [0, n]
    elif len(groups) and isinstance(groups[0], slice):
        off = [groups[0].start or 0]
        for s in groups:
            if (s.start or 0) != off[-1] or s.step not in (None, 1):
                raise ValueError("creature slices must be contiguous and ascending")
            off.append(s.stop)
    else:
        off = [int(x) for x in groups]
    return np.asarray(off, np.int32)


def center_of_mass(position, grid_shape, groups=None, periodic=True):
    """Centre of each node group on the device.  periodic=True:
    swim_base.py:167-177 ``_periodic_com``; False: plain mean (notebook)."""
    lib = _lib.load()
    pos = as_f32(position, position.device)
    off = _offsets(groups, pos.shape[0], pos.device)
    off_d = torch.as_tensor(off, device=pos.device)
    g = off.shape[0] - 1
    out = torch.empty((g, 2), dtype=torch.float32, device=pos.device)
    h, w = grid_shape
    check(lib.mc_center_of_mass(ptr(pos), int(g), ptr(off_d), int(h), int(w), int(bool(periodic)),
                                ptr(out), stream_handle(pos.device)))
    return out


def periodic_com(position, grid_shape):
    """swim_base.py:167-177: centre of all nodes, respecting periodicity."""
    return center_of_mass(position, grid_shape, None, True)[0]


def com_bins(com):
    """utils.py:25-27 index_from_position of centres: (iy, ix) int64."""
    return torch.flip(torch.floor(com).to(torch.int64), dims=(-1,)).T


def rollout_fitness(start_com, end_com, grid_shape):
    """01_es_locomotion.ipynb rollout_fitness: sum over creatures of the x
    component of displacement(end_com, start_com)."""
    from .utils import displacement
    return displacement(grid_shape, end_com, start_com)[:, 0].sum()


class DisplacementTracker:
    """swim_base.py:179-209 on the device: ``init`` (_init_displacement),
    ``update`` (_update_displacement, returns |cumulative|) and
    ``check_stability`` (_check_stability).

    Attached to a ``Rollout`` (``Rollout.track``), the per-step update is
    enqueued by every step the rollout runs; ``update`` is then only needed
    for stepping done outside the rollout.
    """

    def __init__(self, plan, grid_shape, groups=None, num_nodes=None, start_step=0):
        self.plan = plan
        self.grid_shape = tuple(grid_shape)
        n = plan.num_nodes if num_nodes is None else num_nodes
        self.offsets = _offsets(groups, n, None)
        self.n_groups = self.offsets.shape[0] - 1
        self.start_step = int(start_step)
        lib = _lib.load()
        check(lib.mc_track_enable(plan.handle, int(self.n_groups),
                                  self.offsets.ctypes.data_as(ctypes.c_void_p),
                                  self.start_step))

    def init(self, position):
        lib = _lib.load()
        pos = as_f32(position, position.device)
        check(lib.mc_track_reset(self.plan.handle, ptr(pos), stream_handle(pos.device)))
        self.device = pos.device

    def cumulative(self):
        """(n_groups, 2) cumulative minimum-image displacement of the centres."""
        lib = _lib.load()
        out = torch.empty((self.n_groups, 2), dtype=torch.float32, device=self.device)
        check(lib.mc_track_read(self.plan.handle, ptr(out), None, stream_handle(self.device)))
        return out

    def com(self):
        """(n_groups, 2) centre at the last update."""
        lib = _lib.load()
        out = torch.empty((self.n_groups, 2), dtype=torch.float32, device=self.device)
        check(lib.mc_track_read(self.plan.handle, None, ptr(out), stream_handle(self.device)))
        return out

    def displacement(self) -> float:
        """|cumulative| of the first group (the float swim_base returns)."""
        return float(torch.linalg.norm(self.cumulative()[0]))

    @staticmethod
    def check_stability(position, displacement_history):
        """swim_base.py:186-193."""
        if bool(torch.isnan(position).any()):
            raise RuntimeError("problematic displacement found (NaN)")
        if len(displacement_history) >= 100:
            delta = displacement_history[-1] - displacement_history[-100]
            if delta > 100:
                raise RuntimeError(f"problematic displacement found ({delta:.1f} over 100 steps)")
