"""Steric self-avoidance (forces.py:41-118) on the device."""
from __future__ import annotations

import ctypes
from dataclasses import replace

import torch

from . import _lib
from ._lib import McParams, check
from ._runtime import as_f32, plan_for, ptr, stream_handle


def _steric_params(sigma, neighbor_skip, scatter_value) -> McParams:
    p = McParams()
    p.enable_steric = 1
    p.steric_sigma = float(sigma)
    p.steric_neighbor_skip = int(neighbor_skip)
    p.steric_scatter_value = float(scatter_value)
    p.steric_strength = 1.0
    p.cycles_per_step = 0
    return p


def _empty_topology(device):
    return torch.zeros((0, 2), dtype=torch.int32)


def compute_field_steric_force_corrected(nodes, fields, sigma, neighbor_skip, scatter_value=1.0):
    """forces.py:41-81 -- (fields with the new steric field, f_steric (N, 2))."""
    dev = nodes.position.device
    lib = _lib.load()
    pos = as_f32(nodes.position, dev)
    N = pos.shape[0]
    H, W = fields.grid_shape
    z = _empty_topology(dev)
    plan = plan_for(N, z, z, (H, W), (H, W), dev)
    p = _steric_params(sigma, neighbor_skip, scatter_value)
    steric = torch.empty((H, W), device=dev, dtype=torch.float32)
    force = torch.empty((N, 2), device=dev, dtype=torch.float32)
    check(lib.mc_steric_field(plan.handle, ctypes.byref(p), ptr(pos), ptr(steric), ptr(force),
                              None, stream_handle(dev)))
    return replace(fields, steric=steric), force


def update_steric_potential(nodes, fields, sigma=12.5, scatter_value=8.0):
    """forces.py:84-118 -- CIC deposit + Gaussian filter into fields.steric."""
    dev = nodes.position.device
    lib = _lib.load()
    pos = as_f32(nodes.position, dev)
    N = pos.shape[0]
    H, W = fields.grid_shape
    z = _empty_topology(dev)
    plan = plan_for(N, z, z, (H, W), (H, W), dev)
    p = _steric_params(sigma, 0, scatter_value)
    steric = torch.empty((H, W), device=dev, dtype=torch.float32)
    check(lib.mc_steric_field(plan.handle, ctypes.byref(p), ptr(pos), ptr(steric), None, None,
                              stream_handle(dev)))
    return replace(fields, steric=steric)


def deposit_density(position, grid_shape, scatter_value=1.0):
    """The CIC deposit alone (forces.py:95-115), before filtering."""
    dev = position.device
    lib = _lib.load()
    pos = as_f32(position, dev)
    N = pos.shape[0]
    H, W = grid_shape
    z = _empty_topology(dev)
    plan = plan_for(N, z, z, (H, W), (H, W), dev)
    p = _steric_params(1.0, 0, scatter_value)
    steric = torch.empty((H, W), device=dev, dtype=torch.float32)
    dens = torch.empty((H, W), device=dev, dtype=torch.float32)
    check(lib.mc_steric_field(plan.handle, ctypes.byref(p), ptr(pos), ptr(steric), None,
                              ptr(dens), stream_handle(dev)))
    return dens
