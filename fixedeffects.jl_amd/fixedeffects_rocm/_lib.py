"""ctypes binding of libfehip.so (the C-ABI declared in include/fehip.h).

The library is built in-tree (``fixedeffects.jl_amd/lib/libfehip.so``).  There
is no fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("FEHIP_LIB", os.path.join(_PKG_DIR, "lib", "libfehip.so"))

FEHIP_OK = 0
FEHIP_EDIM = -1
FEHIP_EREF = -2
FEHIP_EHIP = -3
FEHIP_EOOM = -4
FEHIP_EARG = -5

FEHIP_FP64 = 0
FEHIP_FP32 = 1
FEHIP_MIXED = 2

EXPORTS = (
    "fehip_solve_residuals_x0",
    "fehip_solve_coefficients_x0",
    "fehip_features",
    "fehip_solve_residuals_learned",
    "fehip_solve_residuals_learned_device",
    "fehip_mlp_create",
    "fehip_mlp_nparams",
    "fehip_mlp_get_params",
    "fehip_mlp_set_params",
    "fehip_mlp_gradient",
    "fehip_mlp_fit",
    "fehip_mlp_predict",
    "fehip_mlp_destroy",
    "fehip_create",
    "fehip_create_i32",
    "fehip_update_weights",
    "fehip_solve_residuals",
    "fehip_solve_residuals_device",
    "fehip_solve_residuals_panels",
    "fehip_solve_residuals_panels_host",
    "fehip_solve_coefficients",
    "fehip_solve_coefficients_full",
    "fehip_normalize",
    "fehip_group",
    "fehip_group_device_noalloc",
    "fehip_set_algorithm",
    "fehip_plan_info",
    "fehip_get_csr",
    "fehip_get_scale",
    "fehip_get_cache",
    "fehip_mul_A",
    "fehip_mul_At",
    "fehip_lsmr_info",
    "fehip_time_kernels",
    "fehip_set_schedule",
    "fehip_autotune",
    "fehip_dims",
    "fehip_last_error",
    "fehip_destroy",
)

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p


class FehipError(RuntimeError):
    """A non-zero status from libfehip, with the library's message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libfehip error {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libfehip.so not found at {LIB_PATH}; build it with `make -C fixedeffects.jl_amd` "
            "or __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.fehip_create.argtypes = [C.POINTER(_vp), C.c_int64, C.c_int32, C.POINTER(_i64p), _i64p, C.POINTER(_dp), _dp,
                               C.c_int32]
    L.fehip_create_i32.argtypes = [C.POINTER(_vp), C.c_int64, C.c_int32, C.POINTER(_i32p), _i64p, C.POINTER(_dp),
                                   _dp, C.c_int32]
    L.fehip_update_weights.argtypes = [_vp, _dp]
    L.fehip_solve_residuals.argtypes = [_vp, C.POINTER(_dp), C.c_int32, C.c_double, C.c_int64, _i64p, _u8p]
    L.fehip_solve_residuals_device.argtypes = [_vp, _vp, C.c_int32, C.c_int64, C.c_double, C.c_int64, _i64p, _u8p]
    L.fehip_solve_residuals_panels.argtypes = [C.POINTER(_vp), C.c_int32, C.POINTER(_vp), C.POINTER(C.c_int32),
                                               C.POINTER(C.c_int64), C.c_double, C.c_int64, C.POINTER(_i64p),
                                               C.POINTER(_u8p)]
    L.fehip_solve_residuals_panels_host.argtypes = [C.POINTER(_vp), C.c_int32, C.POINTER(C.POINTER(_dp)),
                                                    C.POINTER(C.c_int32), C.c_double, C.c_int64, C.POINTER(_i64p),
                                                    C.POINTER(_u8p)]
    L.fehip_solve_coefficients.argtypes = [_vp, _dp, C.c_double, C.c_int64, C.POINTER(_dp), _i64p, _u8p]
    L.fehip_solve_coefficients_full.argtypes = [_vp, _dp, C.c_double, C.c_int64, C.POINTER(_dp), C.POINTER(_dp),
                                                _i64p, _u8p]
    L.fehip_normalize.argtypes = [_vp, C.POINTER(_dp)]
    L.fehip_group.argtypes = [C.c_int64, C.c_int32, C.POINTER(_i64p), _i64p, _i64p]
    L.fehip_group_device_noalloc.argtypes = [C.c_int64, C.c_int32, C.POINTER(_vp), _vp, _i64p]
    L.fehip_set_algorithm.argtypes = [_vp, C.c_int32]
    L.fehip_plan_info.argtypes = [_vp, C.c_int32, C.c_int32, _i64p]
    L.fehip_get_csr.argtypes = [_vp, C.c_int32, _i32p, _i32p]
    L.fehip_get_scale.argtypes = [_vp, C.c_int32, _dp]
    L.fehip_get_cache.argtypes = [_vp, C.c_int32, _dp]
    L.fehip_mul_A.argtypes = [_vp, _dp, _dp, C.c_double, C.c_double]
    L.fehip_mul_At.argtypes = [_vp, _dp, _dp, C.c_double, C.c_double]
    L.fehip_lsmr_info.argtypes = [_vp, C.c_int32, _dp]
    L.fehip_time_kernels.argtypes = [_vp, C.c_int32, C.c_int32, _dp]
    L.fehip_set_schedule.argtypes = [_vp, C.c_int64, C.c_int64, C.c_int32, C.c_int32]
    L.fehip_autotune.argtypes = [_vp, C.c_int32, C.c_int32, _dp]
    L.fehip_dims.argtypes = [_vp, _i64p, _i32p, _i64p]
    L.fehip_last_error.restype = C.c_char_p
    L.fehip_last_error.argtypes = []
    L.fehip_destroy.argtypes = [_vp]
    L.fehip_destroy.restype = None
    _fp = C.POINTER(C.c_float)
    L.fehip_solve_residuals_x0.argtypes = [_vp, C.POINTER(_dp), C.c_int32, C.c_double, C.c_int64, C.POINTER(_dp),
                                           _i64p, _u8p]
    L.fehip_solve_coefficients_x0.argtypes = [_vp, _dp, C.c_double, C.c_int64, _dp, C.POINTER(_dp), _i64p, _u8p]
    L.fehip_features.argtypes = [_vp, _dp, _fp, _dp, _dp]
    L.fehip_solve_residuals_learned.argtypes = [_vp, C.POINTER(_dp), C.c_int32, C.c_double, C.c_int64, _vp, _i64p,
                                                _u8p]
    L.fehip_solve_residuals_learned_device.argtypes = [_vp, _vp, C.c_int32, C.c_int64, C.c_double, C.c_int64, _vp,
                                                       _i64p, _u8p]
    L.fehip_mlp_create.argtypes = [C.POINTER(_vp), C.c_uint64]
    L.fehip_mlp_nparams.argtypes = [_vp, _i64p]
    L.fehip_mlp_get_params.argtypes = [_vp, _fp]
    L.fehip_mlp_set_params.argtypes = [_vp, _fp]
    L.fehip_mlp_gradient.argtypes = [_vp, _fp, _fp, C.c_int64, _fp, _dp]
    L.fehip_mlp_fit.argtypes = [_vp, _fp, _fp, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _dp]
    L.fehip_mlp_predict.argtypes = [_vp, _fp, C.c_int64, _fp]
    L.fehip_mlp_destroy.argtypes = [_vp]
    L.fehip_mlp_destroy.restype = None
    for name in EXPORTS:
        if name not in ("fehip_destroy", "fehip_last_error", "fehip_mlp_destroy"):
            getattr(L, name).restype = C.c_int
    return L


lib = _load()


def check(rc: int) -> None:
    if rc != FEHIP_OK:
        msg = lib.fehip_last_error()
        raise FehipError(rc, msg.decode() if msg else "")
