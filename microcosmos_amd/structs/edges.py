from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass
class Edges:
    """structs/edges.py:6-21 -- edge list (edge -> nodes) and hinge list (hinge -> edges)."""
    pairs: torch.Tensor                # (E, 2) int32 (src, tgt)
    theta: torch.Tensor                # (E,) per-edge material angle
    rest_lengths: torch.Tensor         # (E,)
    bending_pairs: torch.Tensor        # (B, 2) int32 (edge_in, edge_out)
    bending_rest_angles: torch.Tensor  # (B,)
    bending_stiffness: torch.Tensor    # (B,)
