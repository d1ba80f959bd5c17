"""ctypes binding of ``libmicrocosmos_hip.so`` (the C ABI in include/microcosmos_hip.h).

The library is the only compute path.  If it is missing, every entry point
raises: there is no CPU or framework fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int32, c_int64, c_void_p, c_size_t, POINTER
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmicrocosmos_hip.so"

MC_OK = 0
STATUS_NAMES = {0: "MC_OK", 1: "MC_ERR_BAD_SHAPE", 2: "MC_ERR_OOM", 3: "MC_ERR_NON_FINITE",
                4: "MC_ERR_UNSUPPORTED", 5: "MC_ERR_INVALID", 6: "MC_ERR_HIP"}


class MicrocosmosError(RuntimeError):
    """A C-ABI call returned a non-zero mc_status."""

    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


class McParams(ctypes.Structure):
    _fields_ = [("solver_type", c_int32), ("cycles_per_step", c_int32),
                ("stiffness_stretch", c_double), ("stiffness_shear", c_double),
                ("damping", c_double), ("enable_fluid", c_int32), ("viscosity", c_double),
                ("max_fluid_velocity", c_double), ("config_dt", c_double),
                ("ibm_iterations", c_int32), ("synthetic_node_width", c_double),
                ("ibm_kernel_size", c_int32), ("ibm_relaxation", c_double),
                ("node_mass", c_double), ("lambda_trt", c_double),
                ("enable_force_clamp", c_int32), ("enable_steric", c_int32),
                ("steric_strength", c_double), ("steric_sigma", c_double),
                ("steric_neighbor_skip", c_int32), ("steric_scatter_value", c_double),
                ("bending_constraint", c_int32)]


class McState(ctypes.Structure):
    _fields_ = [("position", c_void_p), ("velocity", c_void_p), ("debug_vector", c_void_p),
                ("theta", c_void_p), ("steric", c_void_p), ("fluid_velocity", c_void_p),
                ("f_grid", c_void_p)]


class McEdgeParams(ctypes.Structure):
    _fields_ = [("rest_lengths", c_void_p), ("bending_rest_angles", c_void_p),
                ("bending_stiffness", c_void_p)]


class McActuation(ctypes.Structure):
    _fields_ = [("amp", c_void_p), ("omega", c_void_p), ("phase", c_void_p), ("t0", c_int64)]


class McRecord(ctypes.Structure):
    _fields_ = [("position", c_void_p), ("velocity", c_void_p), ("debug_vector", c_void_p),
                ("steric", c_void_p), ("fluid_velocity", c_void_p), ("stride", c_int64)]


class McActions(ctypes.Structure):
    _fields_ = [("d_rest_length", c_void_p), ("d_bending_angle", c_void_p), ("num_rows", c_int64)]


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "mc_version": (ctypes.c_char_p, []),
    "mc_last_error": (ctypes.c_char_p, []),
    "mc_plan_create": (c_int32, [POINTER(c_void_p), c_int32, c_int32, c_int32, c_int32, c_int32,
                                 c_void_p, c_int32, c_void_p, c_int32, c_int32]),
    "mc_plan_destroy": (c_int32, [c_void_p]),
    "mc_plan_destroy_after": (c_int32, [c_void_p, c_void_p]),
    "mc_plan_status": (c_int32, [c_void_p, c_void_p]),
    "mc_plan_workspace_bytes": (c_int32, [c_void_p, POINTER(c_size_t)]),
    "mc_step": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState), POINTER(McState),
                          POINTER(McEdgeParams), c_double, c_void_p]),
    "mc_rollout": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState), POINTER(McEdgeParams),
                             POINTER(McActuation), c_double, c_int64, c_void_p, c_void_p, c_int32,
                             c_void_p]),
    "mc_profile_step": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                  POINTER(McEdgeParams), POINTER(McActuation), c_double, c_int64,
                                  POINTER(c_double), POINTER(c_double), c_void_p]),
    "mc_profile_fused_step": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                        POINTER(McEdgeParams), POINTER(McActuation), c_double,
                                        c_int64, POINTER(c_double), POINTER(c_double), c_void_p]),
    "mc_gaussian_filter": (c_int32, [c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "mc_steric_field": (c_int32, [c_void_p, POINTER(McParams), c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    "mc_ib_lbm_step": (c_int32, [c_void_p, POINTER(McParams), c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p]),
    "mc_rod_solve": (c_int32, [c_void_p, POINTER(McParams), c_void_p, c_void_p, c_void_p,
                               POINTER(McEdgeParams), c_int32, c_void_p]),
    "mc_batch_plan_create": (c_int32, [POINTER(c_void_p), c_int32, c_void_p, c_void_p, c_void_p,
                                       c_int32, c_void_p, c_int32, c_int32]),
    "mc_plan_instances": (c_int32, [c_void_p, POINTER(c_int32), c_void_p]),
    "mc_batch_rollout": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                   POINTER(McEdgeParams), POINTER(McActuation), c_double, c_int64,
                                   c_void_p, c_void_p]),
    "mc_plan_autotune": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                   POINTER(McEdgeParams), c_double, c_int32, c_void_p]),
    "mc_plan_tuning": (c_int32, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    "mc_plan_set_tuning": (c_int32, [c_void_p, c_int32, c_int32]),
    "mc_rollout_record": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                    POINTER(McEdgeParams), POINTER(McActuation), c_double,
                                    c_int64, POINTER(McRecord), c_int32, c_void_p]),
    "mc_rollout_actions": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                     POINTER(McEdgeParams), POINTER(McActions), c_double, c_int64,
                                     c_void_p, c_void_p]),
    "mc_batch_advance": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                   POINTER(McEdgeParams), POINTER(McActuation), c_void_p,
                                   c_double, c_int64, c_void_p, c_void_p]),
    "mc_rod_adjoint_forward": (c_int32, [c_void_p, POINTER(McParams), POINTER(McState),
                                         POINTER(McEdgeParams), c_double, c_int64, c_int32,
                                         c_void_p]),
    "mc_rod_adjoint_backward": (c_int32, [c_void_p, POINTER(McParams), POINTER(McEdgeParams),
                                          c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p]),
    "mc_rod_adjoint_state": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                       c_void_p]),
    "mc_rod_adjoint_backward_frames": (c_int32, [c_void_p, POINTER(McParams),
                                                 POINTER(McEdgeParams), c_double, c_int32,
                                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_void_p]),
    "mc_rod_adjoint_backward_params": (c_int32, [c_void_p, POINTER(McParams),
                                                 POINTER(McEdgeParams), c_double, c_int32,
                                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "mc_center_of_mass": (c_int32, [c_void_p, c_int32, c_void_p, c_int32, c_int32, c_int32,
                                    c_void_p, c_void_p]),
    "mc_track_enable": (c_int32, [c_void_p, c_int32, c_void_p, c_int32]),
    "mc_track_reset": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "mc_track_read": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "mc_filter_taps": (c_int32, [c_int32, c_double, c_void_p, c_void_p, c_int32]),
    "mc_math_f32": (c_int32, [c_int32, c_void_p, c_int64, c_void_p, c_void_p]),
    "mc_resize_taps": (c_int32, [c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32]),
    "mc_point_loss_workspace": (c_size_t, [c_int64, c_int64]),
    "mc_point_loss": (c_int32, [c_int32, c_void_p, c_int64, c_void_p, c_int64, c_double, c_void_p,
                                c_void_p, c_void_p, c_void_p]),
    "mc_point_loss_static": (c_int32, [c_int32, c_void_p, c_int64, c_void_p, c_int64, c_double,
                                       c_void_p, c_void_p, c_void_p]),
    "mc_frame_stability": (c_int32, [c_void_p, c_int32, c_int64, c_void_p, c_void_p, c_int32,
                                     c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
}

_lib = None


def lib_path() -> Path:
    return _LIB_PATH


def load():
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(
            f"{_LIB_PATH} is missing: build it first (python -m microcosmos_amd.build or "
            "__graft_entry__.build()).  There is no fallback compute path.")
    lib = ctypes.CDLL(str(_LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int):
    if status != MC_OK:
        msg = load().mc_last_error().decode(errors="replace")
        raise MicrocosmosError(status, msg)
