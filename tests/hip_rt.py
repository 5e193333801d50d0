"""Minimal ctypes access to the HIP runtime that libmlmq.so itself links
(device allocation and copies for tests of the device-pointer entry points)."""
import ctypes
import ctypes.util
import os

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        path = "/opt/rocm/lib/libamdhip64.so"
        if not os.path.exists(path):
            path = ctypes.util.find_library("amdhip64") or "libamdhip64.so"
        L = ctypes.CDLL(path)
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _hip = L
    return _hip


class DeviceArray:
    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), max(1, self.nbytes)) == 0
        assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def mem_get_info() -> tuple[int, int]:
    """(free, total) device bytes of the current device."""
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    L = hip()
    L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    assert L.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value, total.value


def to_host(arr: DeviceArray, dtype, count: int) -> np.ndarray:
    out = np.empty(count, dtype=dtype)
    assert hip().hipMemcpy(out.ctypes.data, arr.ptr, out.nbytes, D2H) == 0
    return out
