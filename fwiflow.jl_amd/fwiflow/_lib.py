"""ctypes binding of libfwiflow_hip.so (the C ABI in include/fwiflow.h).

There is no CPU fallback: if the library is missing, loading raises; if no GPU
is present, the compute entry points return FWIFLOW_ERR_HIP and the wrappers
raise ``FwiflowError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FWIFLOW_LIBRARY names another build of the library (e.g. an experimental one)
LIB_PATH = os.environ.get("FWIFLOW_LIBRARY") or os.path.join(os.path.dirname(_HERE), "lib",
                                                               "libfwiflow_hip.so")

STATUS = {
    0: "FWIFLOW_OK", 1: "FWIFLOW_ERR_CALC_ID", 2: "FWIFLOW_ERR_IO", 3: "FWIFLOW_ERR_JSON",
    4: "FWIFLOW_ERR_COURANT", 5: "FWIFLOW_ERR_SHAPE", 6: "FWIFLOW_ERR_UNSUPPORTED",
    7: "FWIFLOW_ERR_HIP", 8: "FWIFLOW_ERR_ARG", 9: "FWIFLOW_ERR_MEMORY",
}
STATUS_CODE = {v: k for k, v in STATUS.items()}

# every symbol include/fwiflow.h declares
EXPORTS = [
    "fwiflow_cufd", "fwiflow_cufd_noalloc", "fwiflow_check", "fwiflow_host_tables", "fwiflow_last_error",
    "fwiflow_version", "fwiflow_plan_create", "fwiflow_plan_destroy", "fwiflow_plan_set_model",
    "fwiflow_plan_set_model_zfast", "fwiflow_plan_set_stf", "fwiflow_plan_set_obs",
    "fwiflow_plan_load_obs", "fwiflow_plan_run", "fwiflow_plan_misfit",
    "fwiflow_plan_gradients", "fwiflow_plan_data", "fwiflow_plan_write_data",
    "fwiflow_plan_residual", "fwiflow_plan_source",
    "fwiflow_plan_dims", "fwiflow_plan_set_checkpoint", "fwiflow_plan_set_adjoint",
    "fwiflow_plan_set_timing",
    "fwiflow_plan_timing", "fwiflow_plan_device_bytes", "fwiflow_plan_shots_per_launch",
    "fwiflow_plan_set_illumination", "fwiflow_plan_illumination",
    "fwiflow_plan_launch_groups", "fwiflow_plan_set_memory_budget", "fwiflow_plan_memory_plan",
    "fwiflow_plan_set_launch", "fwiflow_plan_launch",
    "fwiflow_job_create", "fwiflow_job_destroy", "fwiflow_job_add", "fwiflow_job_run",
    "fwiflow_job_fused",
    "fwiflow_cnn_create", "fwiflow_cnn_destroy", "fwiflow_cnn_num_params",
    "fwiflow_cnn_get_params", "fwiflow_cnn_set_params", "fwiflow_cnn_forward",
    "fwiflow_cnn_gradients", "fwiflow_cnn_train_step",
]


class FwiflowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code
        self.status = STATUS.get(code, str(code))


_lib = None
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


def load():
    """Load the library (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} not found: build it with `python fwiflow.jl_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    L.fwiflow_cufd.argtypes = [dp, dp, dp, dp, dp, dp, dp, dp, dp, C.c_int, C.c_int, C.c_int, ip,
                               C.c_char_p]
    L.fwiflow_cufd_noalloc.argtypes = L.fwiflow_cufd.argtypes
    L.fwiflow_check.argtypes = [C.c_int, C.c_int, ip, C.c_char_p, dp, dp, dp]
    L.fwiflow_host_tables.argtypes = [C.c_char_p, fp, fp, fp, fp]
    L.fwiflow_last_error.restype = C.c_char_p
    L.fwiflow_version.restype = C.c_char_p
    L.fwiflow_plan_create.argtypes = [C.POINTER(vp), C.c_char_p, C.c_int, C.c_int, ip]
    L.fwiflow_plan_destroy.argtypes = [vp]
    L.fwiflow_plan_set_model.argtypes = [vp, dp, dp, dp]
    L.fwiflow_plan_set_model_zfast.argtypes = [vp, dp, dp, dp]
    L.fwiflow_plan_set_stf.argtypes = [vp, dp]
    L.fwiflow_plan_set_obs.argtypes = [vp, fp]
    L.fwiflow_plan_load_obs.argtypes = [vp]
    L.fwiflow_plan_run.argtypes = [vp, C.c_int]
    L.fwiflow_plan_misfit.argtypes = [vp, dp]
    L.fwiflow_plan_gradients.argtypes = [vp, dp, dp, dp, dp]
    L.fwiflow_plan_data.argtypes = [vp, fp]
    L.fwiflow_plan_write_data.argtypes = [vp]
    L.fwiflow_plan_residual.argtypes = [vp, fp]
    L.fwiflow_plan_source.argtypes = [vp, fp]
    L.fwiflow_plan_dims.argtypes = [vp, ip]
    L.fwiflow_plan_set_checkpoint.argtypes = [vp, C.c_int]
    L.fwiflow_plan_set_adjoint.argtypes = [vp, C.c_int]
    L.fwiflow_plan_set_timing.argtypes = [vp, C.c_int]
    L.fwiflow_plan_timing.argtypes = [vp, C.c_char_p, dp, C.POINTER(C.c_longlong)]
    L.fwiflow_plan_device_bytes.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.fwiflow_plan_shots_per_launch.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fwiflow_plan_set_illumination.argtypes = [vp, C.c_int]
    L.fwiflow_plan_illumination.argtypes = [vp, dp]
    L.fwiflow_plan_launch_groups.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fwiflow_plan_set_memory_budget.argtypes = [vp, C.c_double]
    L.fwiflow_plan_memory_plan.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.fwiflow_plan_set_launch.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.fwiflow_plan_launch.argtypes = [vp, ip, ip, ip]
    L.fwiflow_job_create.argtypes = [C.POINTER(vp)]
    L.fwiflow_job_destroy.argtypes = [vp]
    L.fwiflow_job_add.argtypes = [vp, vp]
    L.fwiflow_job_run.argtypes = [vp, C.c_int]
    L.fwiflow_job_fused.argtypes = [vp, ip]
    L.fwiflow_cnn_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_uint]
    L.fwiflow_cnn_destroy.argtypes = [vp]
    L.fwiflow_cnn_num_params.argtypes = [vp, ip]
    L.fwiflow_cnn_get_params.argtypes = [vp, fp]
    L.fwiflow_cnn_set_params.argtypes = [vp, fp]
    L.fwiflow_cnn_forward.argtypes = [vp, fp, fp]
    L.fwiflow_cnn_gradients.argtypes = [vp, fp, fp, fp, dp, fp]
    L.fwiflow_cnn_train_step.argtypes = [vp, fp, fp, fp, C.c_float, dp]
    for name in EXPORTS:
        if name not in ("fwiflow_last_error", "fwiflow_version"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load().fwiflow_last_error().decode(errors="replace")
        raise FwiflowError(rc, msg)


def ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(np.atleast_1d(a), dtype=np.int32)


def version():
    return load().fwiflow_version().decode()
