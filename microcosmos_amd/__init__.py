"""MI355X-native (gfx950) coupled fluid-filament step of Microcosmos.

Drop-in for the reference's step API (simulate.py:15-100): the same pytrees,
ConstraintSolverConfig presets and function signatures, with the whole step
executed by hand-written HIP kernels behind a C ABI (include/microcosmos_hip.h).
"""
from .simulate import Rollout, simulate, step
from .solver.config import (COSSERAT_SCHEME, COSSERAT_SCHEME_NO_FLUID, PBD_SCHEME,
                            PBD_SCHEME_NO_FLUID, ConstraintSolverConfig)
from .solver.pbd import pbd_rod_constraint
from .solver.stable_cosserat import stable_cosserat_constraint
from .structs import Edges, Fields, Nodes

__all__ = ["step", "simulate", "Rollout", "Nodes", "Edges", "Fields", "ConstraintSolverConfig",
           "PBD_SCHEME", "PBD_SCHEME_NO_FLUID", "COSSERAT_SCHEME", "COSSERAT_SCHEME_NO_FLUID",
           "pbd_rod_constraint", "stable_cosserat_constraint"]
