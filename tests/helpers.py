"""Shared test helpers: build the oracle state and the device state from one
seeded workload, run both, compare."""
from __future__ import annotations

from dataclasses import replace

import numpy as np

from oracle import step_ref as R
from microcosmos_amd import workloads as W


def oracle_params(wl, base=R.PBD_SCHEME, **kw):
    return replace(base, **{**wl.config_overrides, **kw})


def oracle_state(wl, dtype=np.float32) -> R.State:
    s = R.State(position=wl.position.copy(), velocity=wl.velocity.copy(),
                debug_vector=np.zeros_like(wl.position), pairs=wl.pairs.copy(),
                theta=wl.theta.copy(), rest_lengths=wl.rest_lengths.copy(),
                bending_pairs=wl.bending_pairs.copy(),
                bending_rest_angles=wl.bending_rest_angles.copy(),
                bending_stiffness=wl.bending_stiffness.copy(), grid_shape=tuple(wl.grid_shape),
                steric=wl.steric.copy(), fluid_velocity=wl.fluid_velocity.copy(),
                f_grid=wl.f_grid.copy())
    return s.astype(dtype) if dtype != np.float32 else s


def run_oracle(wl, steps, params, dtype=np.float32, actuate=True, start=0):
    s = oracle_state(wl, dtype)
    for t in range(start, start + steps):
        if actuate:
            s = replace(s, bending_rest_angles=wl.rest_angles_at(t).astype(dtype))
        s = R.step(s, wl.dt, params)
    return s


def device_state(wl, device="cuda"):
    from microcosmos_amd.workloads import device_state as _ds
    return _ds(wl, device)


def device_config(wl, base=None, **kw):
    from microcosmos_amd import PBD_SCHEME
    base = PBD_SCHEME if base is None else base
    return replace(base, **{**wl.config_overrides, **kw})


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return float(d / n) if n > 0 else float(d)


def pos_rel_l2(grid_shape, a, b):
    """Relative L2 of positions under the minimum image."""
    d = R.displacement(grid_shape, np.asarray(a, np.float64), np.asarray(b, np.float64))
    return float(np.linalg.norm(d) / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-30))


def max_ulp(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7fffffff), a)
    b = np.where(b < 0, -(b & 0x7fffffff), b)
    return int(np.max(np.abs(a - b))) if a.size else 0
