from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch


@dataclass
class Fields:
    """structs/fields.py:5-16 -- grids on the node domain (H, W) and the lattice."""
    grid_shape: tuple               # (H, W), static metadata
    steric: torch.Tensor            # (H, W)
    fluid_velocity: torch.Tensor    # (2, Hf, Wf)
    f_grid: Optional[torch.Tensor]  # (9, Hf, Wf)
    energy: Optional[torch.Tensor] = None

    def __getitem__(self, idx) -> "Fields":
        return Fields(grid_shape=self.grid_shape, steric=self.steric[idx],
                      fluid_velocity=self.fluid_velocity[idx],
                      f_grid=None if self.f_grid is None else self.f_grid[idx],
                      energy=None if self.energy is None else self.energy[idx])
