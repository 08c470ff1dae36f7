"""Minimal HIP runtime access through ctypes: device buffers and copies for
callers that hand libsgnx device pointers.

The runtime is the one libsgnx links (libamdhip64.so.7 from ROCm), resolved
after libsgnx is loaded, so buffers and kernels live in one HIP runtime (a
second runtime in the process, such as the one bundled with a framework wheel,
cannot share device allocations with it).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

H2D, D2H, D2D = 1, 2, 3

_hip = None


def runtime():
    global _hip
    if _hip is None:
        _lib.load()                      # brings libamdhip64.so.7 into the process
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetErrorString.restype = C.c_char_p
        h.hipGetErrorString.argtypes = [C.c_int]
        _hip = h
    return _hip


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}: {runtime().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A float64 device buffer of n elements."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = C.c_void_p()
        _check(runtime().hipMalloc(C.byref(self.ptr), self.n * 8))

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        _check(runtime().hipMemcpy(self.ptr, a.ctypes.data, self.n * 8, H2D))

    def download(self):
        out = np.empty(self.n)
        _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, D2H))
        return out

    def copy_from(self, other):
        assert other.n == self.n
        _check(runtime().hipMemcpy(self.ptr, other.ptr, self.n * 8, D2D))

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _check(runtime().hipDeviceSynchronize())
