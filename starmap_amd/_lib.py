"""ctypes binding of libstarmap_hip.so (include/starmap_hip.h).

There is no fallback: if the library is missing or cannot be loaded, every
entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstarmap_hip.so")

SMP_OK, SMP_E_INVALID, SMP_E_HIP, SMP_E_OOM, SMP_E_STATE, SMP_E_UNSTABLE = range(6)
SMP_FP64, SMP_FP32 = 0, 1
SMP_FILTER_NONE, SMP_FILTER_SUBSTEP, SMP_FILTER_FULL = 0, 1, 2
SMP_COEF_SIGMA_A1, SMP_COEF_SIGMA_T, SMP_COEF_SOURCE, SMP_COEF_FILTER = 0, 1, 2, 3

_ERRNAMES = {SMP_E_INVALID: "invalid", SMP_E_HIP: "hip", SMP_E_OOM: "oom",
             SMP_E_STATE: "state", SMP_E_UNSTABLE: "unstable"}

# every symbol the header declares
EXPORTS = ["smp_last_error", "smp_version", "smp_create", "smp_destroy", "smp_grid_types",
           "smp_component_shape", "smp_set_state", "smp_get_state", "smp_set_coeff",
           "smp_step", "smp_snapshot", "smp_interp", "smp_observables",
           "smp_last_step_timing", "smp_stream", "smp_device_bytes", "smp_plan_stats",
           "smp_set_schedule", "smp_get_schedule", "smp_checkpoint_save", "smp_checkpoint_load",
           "smp_checkpoint_free", "smp_batch_create", "smp_batch_destroy", "smp_batch_run",
           "smp_batch_timing", "smp_adjoint_reset", "smp_adjoint_seed", "smp_adjoint",
           "smp_adjoint_get", "smp_adjoint_timing", "smp_moment_norms", "smp_probe",
           "smp_check_finite", "smp_probe_noalloc", "smp_num_components",
           "smp_adjoint_set_memory", "smp_adjoint_stats", "smp_adjoint_schedule"]


class SmpError(RuntimeError):
    """A status other than SMP_OK returned by the library
    (the MEX gateway maps the same codes to starmap:hip:<code>)."""

    def __init__(self, code, msg):
        super().__init__("starmap:hip:%s: %s" % (_ERRNAMES.get(code, code), msg))
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [("n_sys", ctypes.c_int32),
                ("n", ctypes.c_int32 * 3),
                ("bc", ctypes.c_int32 * 3),
                ("h", ctypes.c_double * 3),
                ("row_ptr", ctypes.c_void_p * 3),
                ("col", ctypes.c_void_p * 3),
                ("val", ctypes.c_void_p * 3),
                ("mom_order", ctypes.c_void_p),
                ("f_order", ctypes.c_void_p),
                ("f_position", ctypes.c_int32),
                ("precision", ctypes.c_int32),
                ("device", ctypes.c_int32),
                ("tile", ctypes.c_int32 * 3),
                ("blocks_per_cu", ctypes.c_int32),
                ("cell_waves", ctypes.c_int32)]


_lib = None


def lib():
    """Load the library (once); raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libstarmap_hip.so is not built (run `python -m starmap_amd.build`)")
    # torch ships its own HIP runtime (libamdhip64.so.7).  Loaded first, the
    # library binds the system runtime under the same soname, and torch then
    # finds no GPU in the process; with torch loaded first both share its
    # runtime, so torch's streams and synchronisation see the library's work.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "smp_last_error": (ctypes.c_char_p, []),
        "smp_version": (ctypes.c_char_p, []),
        "smp_create": (ctypes.c_int, [vp, vp]),
        "smp_destroy": (ctypes.c_int, [vp]),
        "smp_grid_types": (ctypes.c_int, [vp, vp]),
        "smp_component_shape": (ctypes.c_int, [vp, i32, vp]),
        "smp_set_state": (ctypes.c_int, [vp, i32, vp, i64]),
        "smp_get_state": (ctypes.c_int, [vp, i32, vp, i64]),
        "smp_set_coeff": (ctypes.c_int, [vp, i32, vp, i32, i32, vp, dbl, i64]),
        "smp_step": (ctypes.c_int, [vp, dbl, i64]),
        "smp_snapshot": (ctypes.c_int, [vp, vp, i32]),
        "smp_interp": (ctypes.c_int, [vp, dbl, vp, i64]),
        "smp_observables": (ctypes.c_int, [vp, vp, vp, vp]),
        "smp_last_step_timing": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "smp_stream": (vp, [vp]),
        "smp_device_bytes": (i64, [vp]),
        "smp_plan_stats": (ctypes.c_int, [vp, i32, vp, vp, vp, vp]),
        "smp_set_schedule": (ctypes.c_int, [vp, vp, i32, i32]),
        "smp_get_schedule": (ctypes.c_int, [vp, vp, vp, vp]),
        "smp_checkpoint_save": (ctypes.c_int, [vp, i32]),
        "smp_checkpoint_load": (ctypes.c_int, [vp, i32]),
        "smp_checkpoint_free": (ctypes.c_int, [vp, i32]),
        "smp_batch_create": (ctypes.c_int, [vp, i32, vp]),
        "smp_batch_destroy": (ctypes.c_int, [vp]),
        "smp_batch_run": (ctypes.c_int, [vp, vp, vp, i32]),
        "smp_batch_timing": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp]),
        "smp_adjoint_reset": (ctypes.c_int, [vp]),
        "smp_adjoint_seed": (ctypes.c_int, [vp, i32, vp, i64]),
        "smp_adjoint": (ctypes.c_int, [vp, dbl, i64, i32]),
        "smp_adjoint_get": (ctypes.c_int, [vp, i32, i32, vp, i64]),
        "smp_adjoint_timing": (ctypes.c_int, [vp, vp, vp]),
        "smp_moment_norms": (ctypes.c_int, [vp, vp, vp]),
        "smp_probe": (ctypes.c_int, [vp, vp, vp, vp]),
        "smp_check_finite": (ctypes.c_int, [vp, vp]),
        "smp_probe_noalloc": (ctypes.c_int, [vp, vp, vp, vp]),
        "smp_num_components": (ctypes.c_int, [vp, vp]),
        "smp_adjoint_set_memory": (ctypes.c_int, [vp, i64]),
        "smp_adjoint_stats": (ctypes.c_int, [vp, vp]),
        "smp_adjoint_schedule": (ctypes.c_int, [i64, i32, i64, i32, vp, vp, i64, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class SmpUnstable(SmpError):
    """SMP_E_UNSTABLE: NaN or infinite values in the state, the sign of a
    time step beyond the stability limit (par.cn too large)."""


def check(rc):
    if rc != SMP_OK:
        cls = SmpUnstable if rc == SMP_E_UNSTABLE else SmpError
        raise cls(rc, lib().smp_last_error().decode())


def ptr(a):
    return a.ctypes.data if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
