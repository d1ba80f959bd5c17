"""stable_cosserat_constraint (solver/stable_cosserat.py:13-119) -- one
iteration on the device (theta pass + implicit inertial position update toward
y = nodes.debug_vector)."""
from __future__ import annotations


def stable_cosserat_constraint(iter, carry, config, grid_shape):
    from .._runtime import rod_iterations
    return rod_iterations(carry, config, grid_shape, 1, solver="cosserat")
