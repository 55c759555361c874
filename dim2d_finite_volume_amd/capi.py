"""ctypes binding of the C ABI declared in include/dim2d_hip.h.

The library is the in-tree build `dim2d_finite_volume_amd/lib/libdim2d_hip.so`
(built by `make -C dim2d_finite_volume_amd` or `__graft_entry__.build()`); the
environment variable DIM2D_HIP_LIBRARY selects another build of the same ABI.
There is no fallback: if the library is missing, `load()` raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DIM2D_HIP_LIBRARY") or os.path.join(HERE, "lib", "libdim2d_hip.so")

OK = 0
EINVAL, EDEVICE, ENOMEM, ESTATE, ENONFINITE = -1, -2, -3, -4, -5
BC_REFLECTION, BC_PERIODIC, BC_WALL, BC_OPEN = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
INT_RK3TVD, INT_LSRK3 = 0, 1
OPS_MATTSSON, OPS_COCKBURN_GAU = 0, 1

# every symbol declared in include/dim2d_hip.h
EXPORTS = (
    "dim2d_last_error", "dim2d_version", "dim2d_ctx_create", "dim2d_ctx_destroy",
    "dim2d_upload", "dim2d_download", "dim2d_upload_device", "dim2d_download_device",
    "dim2d_time_dev", "dim2d_apply_bc", "dim2d_rk3_step", "dim2d_integrate", "dim2d_cfl_dt",
    "dim2d_diagnostics", "dim2d_time", "dim2d_synchronize", "dim2d_set_tiles",
    "dim2d_set_stage_fusion", "dim2d_profile_enable", "dim2d_profile_read",
    "dim2d_get_params", "dim2d_openbc_undermined", "dim2d_set_far_field", "dim2d_extend",
    "dim2d_adapt_domain", "dim2d_get_growth", "dim2d_shrink", "dim2d_openbc_undermined_decrease",
    "dim2d_check_steady_state", "dim2d_time_dev_noalloc",
    "dim2d_batch_create", "dim2d_batch_destroy", "dim2d_batch_upload", "dim2d_batch_download",
    "dim2d_batch_integrate", "dim2d_batch_set_tiles", "dim2d_batch_synchronize",
    "dim2d_batch_autotune", "dim2d_batch_set_stage_fusion", "dim2d_batch_get_tiles",
)


class Params(ctypes.Structure):
    """dim2d_params (include/dim2d_hip.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("bc_size", ctypes.c_int32),
        ("bc_w", ctypes.c_int32), ("bc_e", ctypes.c_int32), ("bc_s", ctypes.c_int32),
        ("bc_n", ctypes.c_int32), ("precision", ctypes.c_int32), ("gravity_ac", ctypes.c_int32),
        ("x_min", ctypes.c_double), ("x_max", ctypes.c_double), ("y_min", ctypes.c_double),
        ("y_max", ctypes.c_double), ("re", ctypes.c_double), ("we", ctypes.c_double),
        ("pr", ctypes.c_double), ("viscous_r", ctypes.c_double), ("cv_r", ctypes.c_double),
        ("gravity", ctypes.c_double), ("wall_contact_angle", ctypes.c_double),
        ("wall_heat_flux", ctypes.c_double), ("integrator", ctypes.c_int32),
        ("operators", ctypes.c_int32),
    ]


class Dim2dError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()' or make -C dim2d_finite_volume_amd)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.c_int32
    d = ctypes.c_double
    sig = {
        "dim2d_last_error": (ctypes.c_char_p, []),
        "dim2d_version": (ctypes.c_char_p, []),
        "dim2d_ctx_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(vp)]),
        "dim2d_ctx_destroy": (ctypes.c_int, [vp]),
        "dim2d_upload": (ctypes.c_int, [vp, dp]),
        "dim2d_download": (ctypes.c_int, [vp, dp]),
        "dim2d_upload_device": (ctypes.c_int, [vp, vp]),
        "dim2d_download_device": (ctypes.c_int, [vp, vp]),
        "dim2d_time_dev": (ctypes.c_int, [vp, d, dp]),
        "dim2d_time_dev_noalloc": (ctypes.c_int, [ctypes.POINTER(Params), vp, vp, vp]),
        "dim2d_apply_bc": (ctypes.c_int, [vp, d]),
        "dim2d_rk3_step": (ctypes.c_int, [vp, d, d]),
        "dim2d_integrate": (ctypes.c_int, [vp, d, d, d, i32]),
        "dim2d_cfl_dt": (ctypes.c_int, [vp, d, dp]),
        "dim2d_diagnostics": (ctypes.c_int, [vp, dp]),
        "dim2d_time": (ctypes.c_int, [vp, dp]),
        "dim2d_synchronize": (ctypes.c_int, [vp]),
        "dim2d_set_tiles": (ctypes.c_int, [vp, i32, i32]),
        "dim2d_set_stage_fusion": (ctypes.c_int, [vp, i32, i32]),
        "dim2d_profile_enable": (ctypes.c_int, [vp, i32]),
        "dim2d_profile_read": (ctypes.c_int, [vp, dp]),
        "dim2d_get_params": (ctypes.c_int, [vp, ctypes.POINTER(Params)]),
        "dim2d_check_steady_state": (ctypes.c_int, [vp, d, dp, ctypes.POINTER(i32)]),
        "dim2d_openbc_undermined": (ctypes.c_int, [vp, i32, d, ctypes.POINTER(i32)]),
        "dim2d_set_far_field": (ctypes.c_int, [vp, ctypes.c_void_p]),
        "dim2d_extend": (ctypes.c_int, [vp, i32, i32, i32, i32]),
        "dim2d_adapt_domain": (ctypes.c_int, [vp, i32, i32, d, i32, ctypes.POINTER(i32)]),
        "dim2d_get_growth": (ctypes.c_int, [vp, ctypes.POINTER(i32)]),
        "dim2d_shrink": (ctypes.c_int, [vp, i32, i32, i32, i32]),
        "dim2d_openbc_undermined_decrease": (ctypes.c_int, [vp, i32, d, ctypes.POINTER(i32)]),
        "dim2d_batch_create": (ctypes.c_int, [ctypes.POINTER(Params), i32, ctypes.POINTER(vp)]),
        "dim2d_batch_destroy": (ctypes.c_int, [vp]),
        "dim2d_batch_upload": (ctypes.c_int, [vp, i32, dp]),
        "dim2d_batch_download": (ctypes.c_int, [vp, i32, dp]),
        "dim2d_batch_integrate": (ctypes.c_int, [vp, dp, i32]),
        "dim2d_batch_set_tiles": (ctypes.c_int, [vp, i32, i32, i32]),
        "dim2d_batch_synchronize": (ctypes.c_int, [vp]),
        "dim2d_batch_autotune": (ctypes.c_int, [vp, dp, ctypes.c_int32, dp]),
        "dim2d_batch_set_stage_fusion": (ctypes.c_int, [vp, i32, i32, i32]),
        "dim2d_batch_get_tiles": (ctypes.c_int, [vp, ctypes.POINTER(i32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, where: str) -> None:
    if status != OK:
        msg = load().dim2d_last_error().decode(errors="replace")
        raise Dim2dError(status, where, msg)
