"""Device buffers for callers that keep columns resident in HBM.

Uses the HIP runtime that libfehip.so itself links (/opt/rocm's
libamdhip64.so.7, resolved by soname after libfehip is loaded), so buffers and
the solver share one runtime.  Do not import torch in the same process: the
torch wheel ships its own, separately loaded HIP/HSA runtime, and two HSA
runtimes in one process do not share the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib as _fehip  # noqa: F401  (loads libamdhip64.so.7 first)

_hip = C.CDLL("libamdhip64.so.7")
_hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.argtypes = []
_hip.hipGetErrorString.restype = C.c_char_p
_hip.hipGetErrorString.argtypes = [C.c_int]

H2D, D2H, D2D = 1, 2, 3


def _chk(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {_hip.hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = _hip.hipGetDeviceCount(C.byref(n))
    return int(n.value) if rc == 0 else 0


def synchronize():
    _chk(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd block of float64s."""

    def __init__(self, count: int):
        self.count = int(count)
        self.ptr = C.c_void_p()
        _chk(_hip.hipMalloc(C.byref(self.ptr), max(1, self.count) * 8), "hipMalloc")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size <= self.count
        _chk(_hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.size * 8, H2D), "hipMemcpy H2D")

    def download(self, count: int = None) -> np.ndarray:
        n = self.count if count is None else int(count)
        out = np.empty(n, np.float64)
        _chk(_hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, n * 8, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceBuffer"):
        assert other.count <= self.count
        _chk(_hip.hipMemcpy(self.ptr, other.ptr, other.count * 8, D2D), "hipMemcpy D2D")

    def free(self):
        if self.ptr.value:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
