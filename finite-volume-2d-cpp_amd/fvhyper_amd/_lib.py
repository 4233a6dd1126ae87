"""ctypes binding of libfvh.so (include/fvh.h).

The library must be built (``python finite-volume-2d-cpp_amd/build.py``);
loading fails loudly otherwise -- there is no CPU fallback in the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfvh.so")

FVH_MAX_ALPHA = 8
FVH_MAX_BC = 16
FVH_NAME_LEN = 64

FLUX = {"roe_split": 0, "roe_joined": 1, "roe_viscous": 2}
PHI = {"zero": 0, "michalak": 1, "venkat": 2, "minmod": 3}
DT = {"fixed": 0, "sum_cfl": 1, "edge_min": 2, "edge_min_viscous": 3, "sum_viscous": 4}
BC = {"copy": 0, "mirror": 1, "inlet_outlet": 2, "farfield": 3, "noslip": 4, "farfield_v": 5}
FLAG_GRADIENTS, FLAG_LIMITERS, FLAG_LINEAR, FLAG_GLOBAL_DT, FLAG_POSITIVITY, FLAG_SMOOTH = 1, 2, 4, 8, 16, 32
FLAG_DIFFUSIVE = 64
FLAG_SMOOTH_KINKS = 128
RK5 = [0.0695, 0.1602, 0.2898, 0.506, 1.0]


class BcSpec(C.Structure):
    _fields_ = [("name", C.c_char * FVH_NAME_LEN), ("kind", C.c_int32),
                ("rho", C.c_double), ("u", C.c_double), ("v", C.c_double), ("p", C.c_double)]


class Config(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("flux", C.c_int32), ("gamma", C.c_double),
        ("phi", C.c_int32), ("phi_yt", C.c_double), ("K", C.c_double),
        ("dt_kind", C.c_int32), ("dt_fixed", C.c_double), ("cfl", C.c_double),
        ("flags", C.c_uint32), ("n_alpha", C.c_int32), ("alpha", C.c_double * FVH_MAX_ALPHA),
        ("n_bc", C.c_int32), ("bc", BcSpec * FVH_MAX_BC),
        ("renumber", C.c_int32), ("flux_mode", C.c_int32), ("tile_cells", C.c_int32),
        ("mu", C.c_double), ("pr", C.c_double), ("cp", C.c_double), ("r_gas", C.c_double),
        ("stage_threads", C.c_int32), ("stage_launch", C.c_int32),
        ("kink_width", C.c_double), ("kink_scale", C.c_double * 5),
    ]


class RunOptions(C.Structure):
    _fields_ = [("max_time", C.c_double), ("max_step", C.c_uint32),
                ("print_interval", C.c_uint32), ("tolerance", C.c_double),
                ("stop_var", C.c_int32), ("reserved", C.c_int32)]


class AdjointStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("snaps", C.c_int32), ("forward_steps", C.c_int64),
                ("checkpoint_bytes", C.c_double), ("work_bytes", C.c_double),
                ("forward_ms", C.c_double), ("total_ms", C.c_double), ("recorded_steps", C.c_int32),
                ("reserved", C.c_int32)]


class Status(C.Structure):
    _fields_ = [("step", C.c_uint32), ("stopped", C.c_uint32), ("time", C.c_double),
                ("time_prev", C.c_double), ("R", C.c_double * 4), ("R0", C.c_double * 4)]


ROW_CB = C.CFUNCTYPE(None, C.c_uint32, C.POINTER(Status), C.c_void_p)

# every symbol include/fvh.h declares
EXPORTS = [
    "fvh_last_error", "fvh_abi_version", "fvh_mesh_read_msh", "fvh_mesh_read_su2", "fvh_mesh_from_arrays",
    "fvh_mesh_free", "fvh_mesh_counts", "fvh_mesh_export", "fvh_mesh_name",
    "fvh_config_default", "fvh_solver_create", "fvh_solver_free", "fvh_solver_set_state",
    "fvh_solver_get_state", "fvh_solver_save", "fvh_solver_load", "fvh_solver_apply_bounds", "fvh_solver_phases", "fvh_solver_run",
    "fvh_solver_advance", "fvh_solver_run_begin", "fvh_solver_run_continue", "fvh_solver_run_log",
    "fvh_stream_copy", "fvh_solver_synchronize", "fvh_solver_status", "fvh_solver_profile",
    "fvh_solver_kernel_times", "fvh_solver_kernel_name", "fvh_solver_stage_bytes", "fvh_solver_tile_cells",
    "fvh_solver_use_graphs", "fvh_batch_create", "fvh_batch_size", "fvh_batch_instance",
    "fvh_batch_status", "fvh_batch_sums", "fvh_autotune", "fvh_objective", "fvh_adjoint",
    "fvh_phase_gradients", "fvh_phase_limiters", "fvh_phase_time_derivatives", "fvh_phase_update_cells",
    "fvh_phase_update_bounds", "fvh_phase_min_dt", "fvh_phase_residuals", "fvh_solver_stage_info",
    "fvh_min_dt_device",
]

_lib = None


class FvhError(RuntimeError):
    pass


def _share_hip_runtime_with_torch():
    """Make libfvh.so bind the HIP runtime torch ships, when torch is there.

    libfvh.so needs libamdhip64.so.7 (found through its run path in the ROCm
    install).  torch bundles its own libamdhip64 with the same soname.  Two
    copies of the runtime in one process leave the one loaded second without
    devices ("No HIP GPUs are available"), so whichever library comes first
    decides.  Loading torch's copy by path, with global symbols, before
    libfvh.so makes both use it in either import order, without the cost of
    importing torch.  Without torch, libfvh.so takes the ROCm runtime."""
    import importlib.util
    import sys
    if "torch" in sys.modules:
        return   # torch's runtime is already in the process
    try:
        spec = importlib.util.find_spec("torch")
    except Exception:   # a broken torch install: libfvh.so takes the ROCm runtime
        return
    for d in (spec.submodule_search_locations or []) if spec else []:
        p = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(p):
            try:
                C.CDLL(p, mode=C.RTLD_GLOBAL)
            except OSError:
                pass
            return


def load(path: str = LIB_PATH):
    """Load libfvh.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FvhError("libfvh.so is not built at %s (run finite-volume-2d-cpp_amd/build.py)" % path)
    _share_hip_runtime_with_torch()
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.fvh_last_error.restype = C.c_char_p
    L.fvh_abi_version.restype = C.c_int32
    L.fvh_mesh_read_msh.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.fvh_mesh_read_su2.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.fvh_mesh_from_arrays.argtypes = [C.c_int32, dp, dp, C.c_int32, ip, C.c_int32, ip, ip, C.c_int32,
                                       C.POINTER(C.c_char_p), C.POINTER(vp)]
    L.fvh_mesh_free.argtypes = [vp]
    L.fvh_mesh_free.restype = None
    L.fvh_mesh_counts.argtypes = [vp, C.POINTER(C.c_int64)]
    L.fvh_mesh_export.argtypes = [vp, C.c_char_p, vp, C.c_int64]
    L.fvh_mesh_export.restype = C.c_int64
    L.fvh_mesh_name.argtypes = [vp, C.c_int32, C.c_char_p, C.c_int32]
    L.fvh_config_default.argtypes = [C.POINTER(Config)]
    L.fvh_config_default.restype = None
    L.fvh_solver_create.argtypes = [vp, C.POINTER(Config), C.POINTER(vp)]
    L.fvh_solver_free.argtypes = [vp]
    L.fvh_solver_free.restype = None
    L.fvh_solver_set_state.argtypes = [vp, dp]
    L.fvh_solver_get_state.argtypes = [vp, dp]
    L.fvh_solver_save.argtypes = [vp, C.c_char_p]
    L.fvh_solver_load.argtypes = [vp, C.c_char_p]
    L.fvh_solver_apply_bounds.argtypes = [vp]
    L.fvh_solver_phases.argtypes = [vp, C.c_double, dp, dp, dp, dp, dp, dp]
    L.fvh_solver_run.argtypes = [vp, C.POINTER(RunOptions), C.POINTER(Status)]
    L.fvh_solver_advance.argtypes = [vp, C.c_int32]
    L.fvh_solver_run_begin.argtypes = [vp]
    L.fvh_solver_run_continue.argtypes = [vp, C.POINTER(RunOptions), C.c_uint32, C.POINTER(Status)]
    L.fvh_solver_run_log.argtypes = [vp, C.POINTER(RunOptions), ROW_CB, vp, C.POINTER(Status)]
    L.fvh_stream_copy.argtypes = [C.c_double, C.c_int32, dp]
    L.fvh_solver_synchronize.argtypes = [vp]
    L.fvh_solver_status.argtypes = [vp, C.POINTER(Status)]
    L.fvh_solver_profile.argtypes = [vp, C.c_int32]
    L.fvh_solver_kernel_times.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), dp]
    L.fvh_solver_kernel_name.argtypes = [C.c_int32]
    L.fvh_solver_kernel_name.restype = C.c_char_p
    L.fvh_solver_stage_bytes.argtypes = [vp]
    L.fvh_solver_stage_bytes.restype = C.c_double
    L.fvh_solver_tile_cells.argtypes = [vp]
    L.fvh_solver_stage_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.fvh_solver_use_graphs.argtypes = [vp, C.c_int32]
    L.fvh_batch_create.argtypes = [C.c_int32, C.POINTER(vp), C.POINTER(Config), C.POINTER(vp)]
    L.fvh_batch_size.argtypes = [vp]
    L.fvh_batch_instance.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64)]
    L.fvh_batch_status.argtypes = [vp, C.c_int32, C.POINTER(Status)]
    L.fvh_batch_sums.argtypes = [vp, dp]
    L.fvh_objective.argtypes = [vp, dp, C.c_int32, dp, dp, dp]
    L.fvh_adjoint.argtypes = [vp, dp, C.c_int32, dp, dp, C.c_int32, C.c_double, dp, dp, dp,
                              C.POINTER(AdjointStats)]
    L.fvh_autotune.argtypes = [C.c_int32, C.POINTER(vp), C.POINTER(Config), dp, C.c_int32,
                               C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp, C.c_int32]
    L.fvh_phase_gradients.argtypes = [vp, dp, dp, dp]
    L.fvh_phase_limiters.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.fvh_phase_time_derivatives.argtypes = [vp, dp, dp, dp, dp, dp]
    L.fvh_phase_update_cells.argtypes = [vp, dp, dp, dp, dp, C.c_double, C.c_int64]
    L.fvh_phase_update_bounds.argtypes = [vp, dp, dp, dp, dp]
    L.fvh_phase_min_dt.argtypes = [vp, dp, C.c_int64]
    L.fvh_phase_residuals.argtypes = [vp, dp, dp]
    L.fvh_min_dt_device.argtypes = [vp, C.c_int64, vp]
    for name in EXPORTS:
        getattr(L, name)
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        raise FvhError("fvh error %d: %s" % (rc, load().fvh_last_error().decode(errors="replace")))
    return rc


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))
