from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass
class Nodes:
    """structs/nodes.py:5-15 -- per-node state, float32 (N, 2) arrays in (x, y)."""
    position: torch.Tensor      # (N, 2)
    velocity: torch.Tensor      # (N, 2)
    color_bend: torch.Tensor    # (N,) rendering only, carried through
    debug_vector: torch.Tensor  # (N, 2) Cosserat inertial target y (or zeros)

    def __getitem__(self, idx) -> "Nodes":
        return Nodes(position=self.position[idx], velocity=self.velocity[idx],
                     color_bend=self.color_bend[idx], debug_vector=self.debug_vector[idx])

    def __len__(self) -> int:
        return len(self.position)
