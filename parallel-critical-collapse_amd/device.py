"""Device arrays on the HIP runtime that libcc_gpu.so itself links (no torch).

Only plumbing: allocation and copies for callers that keep inputs resident
in HBM and call the *_device entry points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from . import load
        load()                                   # libcc_gpu.so pulls in libamdhip64.so.7
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceProperties.argtypes = [C.c_void_p, C.c_int]
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        rc = hip().hipMemcpy(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p), self.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")
