"""pbd_rod_constraint (solver/pbd.py:14-91) -- one Jacobi iteration on the device.

Inside ``step`` the constraint loop runs fused in the rod kernel; this
callable is the reference's plug-in signature for callers that iterate it
themselves (``fori_loop(0, n, fn, (nodes, edges))``).
"""
from __future__ import annotations

from dataclasses import replace


def pbd_rod_constraint(iter, carry, config, grid_shape):
    from .._runtime import rod_iterations
    return rod_iterations(carry, config, grid_shape, 1, solver="pbd")
