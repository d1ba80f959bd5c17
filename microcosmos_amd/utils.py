"""Geometry and filter primitives (utils.py:9-49).

``gaussian_filter`` runs the device filter; the small geometry helpers are
elementwise torch expressions of the reference formulas, for callers that set
up states (they are not on the step path, which does this arithmetic inside
the HIP kernels).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from ._runtime import as_f32, plan_for, ptr, stream_handle

EPSIL = float(np.finfo(np.float32).eps)  # utils.py:9


def _size_xy(size, like):
    s = torch.as_tensor(size, dtype=like.dtype, device=like.device)
    return torch.flip(s, dims=(0,))


def displacement(size, x, y):
    """utils.py:18-23 -- minimum-image mod(x - y + S/2, S) - S/2 with S = (W, H)."""
    s = _size_xy(size, x)
    return torch.remainder(x - y + s / 2.0, s) - s / 2.0


def periodic_boundary(size, position):
    """utils.py:29-33."""
    return torch.remainder(position, _size_xy(size, position))


def index_from_position(position):
    """utils.py:25-27 -- (iy, ix)."""
    return torch.flip(torch.floor(position).to(torch.int64), dims=(-1,)).T


def norm(v, axis=-1, keepdims=True):
    """utils.py:12-15."""
    return torch.clamp(torch.linalg.norm(v + EPSIL, dim=axis, keepdim=keepdims), min=EPSIL)


def gaussian_filter(data, sigma):
    """utils.py:35-49 -- periodic Gaussian filter on the device: the direct
    separable convolution when the periodic kernel is short, else the LDS
    FFT."""
    dev = data.device
    lib = _lib.load()
    d = as_f32(data, dev)
    H, W = d.shape
    z = torch.zeros((0, 2), dtype=torch.int32)
    plan = plan_for(0, z, z, (H, W), (H, W), dev)
    out = torch.empty_like(d)
    check(lib.mc_gaussian_filter(plan.handle, ptr(d), ptr(out), float(sigma), stream_handle(dev)))
    return out
