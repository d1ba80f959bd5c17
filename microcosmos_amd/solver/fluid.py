"""immersed_boundary_interaction (solver/fluid.py:202-386) on the device:
sparse IB multi-direct forcing + fused TRT/Guo collide-stream + node drag."""
from __future__ import annotations

import ctypes
from dataclasses import replace

import torch

from .. import _lib
from .._lib import check
from .._runtime import as_f32, params_from_config, plan_for, ptr, stream_handle


def immersed_boundary_interaction(nodes, edges, fields, f_grid, dt, solver_config):
    """Returns (nodes with v - dv, fields with fluid_velocity = u*, f_new)."""
    dev = f_grid.device
    lib = _lib.load()
    p = params_from_config(solver_config)
    f_in = as_f32(f_grid, dev)
    H, W = fields.grid_shape
    N = nodes.position.shape[0]
    plan = plan_for(N, edges.pairs, edges.bending_pairs, (H, W), tuple(f_in.shape[1:]), dev)
    pos = as_f32(nodes.position, dev)
    vel = as_f32(nodes.velocity, dev)
    o_vel = torch.empty_like(vel)
    o_fv = torch.empty((2,) + tuple(f_in.shape[1:]), device=dev, dtype=torch.float32)
    o_f = torch.empty_like(f_in)
    check(lib.mc_ib_lbm_step(plan.handle, ctypes.byref(p), ptr(pos), ptr(vel), ptr(f_in),
                             ptr(o_vel), ptr(o_fv), ptr(o_f), None, float(dt),
                             stream_handle(dev)))
    return replace(nodes, velocity=o_vel), replace(fields, fluid_velocity=o_fv), o_f
