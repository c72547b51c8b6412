"""Minimal device-memory plumbing over the HIP runtime (hipMalloc / hipMemcpy),
for callers that keep potentials and gradients resident in HBM and use
SotSolver.eval_device."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                   "libamdhip64.so"))
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


_H2D, _D2H = 1, 2


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceArray:
    """A float64 array in device memory."""

    def __init__(self, n: int, device: int = 0):
        self.n = int(n)
        self.ptr = C.c_void_p()
        _ok(hip().hipSetDevice(device), "hipSetDevice")
        _ok(hip().hipMalloc(C.byref(self.ptr), max(1, self.n) * 8), "hipMalloc")

    @classmethod
    def from_host(cls, a, device: int = 0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size, device)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, self.n * 8, _H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, _D2H), "hipMemcpy D2H")
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_d2d(dst: DeviceArray, src: DeviceArray):
    assert dst.n == src.n
    _ok(hip().hipMemcpy(dst.ptr, src.ptr, src.n * 8, 3), "hipMemcpy D2D")


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
