"""Topology and initial-state builders (graph.py:156-401,533-596)."""
from __future__ import annotations

import numpy as np
import torch

from .structs import Fields
from .topology import compute_bending_pairs, line_instances, make_edges  # noqa: F401
from .graph_builder import GraphBuilder, initialize_positions, parse_graph_json  # noqa: F401

_E = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
_W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])


def initialize_f_from_velocity(velocity_field, density=1.0):
    """graph.py:533-540 -- feq(density, u) with float32 lattice weights."""
    u = torch.as_tensor(velocity_field, dtype=torch.float32)
    usq = u[0] * u[0] + u[1] * u[1]
    out = []
    for i in range(9):
        ex, ey = int(_E[i, 0]), int(_E[i, 1])
        eu = ex * u[0] + ey * u[1]
        w = float(np.float32(density * _W[i]))
        out.append(w * (1.0 + 3.0 * eu + 4.5 * (eu * eu) - 1.5 * usq))
    return torch.stack(out)


def make_fields(size, device="cuda"):
    """graph.py:586-596 -- zero velocity, equilibrium lattice, zero steric."""
    h, w = size
    fv = torch.zeros((2, h, w), dtype=torch.float32)
    f = initialize_f_from_velocity(fv)
    return Fields(grid_shape=(h, w), steric=torch.zeros((h, w), dtype=torch.float32, device=device),
                  fluid_velocity=fv.to(device), f_grid=f.to(device))
