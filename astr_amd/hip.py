"""Minimal ctypes access to the HIP runtime for host plumbing: device buffers,
copies and events (used to keep benchmark inputs resident in HBM)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_rt = None


def rt():
    global _rt
    if _rt is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", os.path.join("/opt/rocm/lib", "libamdhip64.so")):
            try:
                _rt = C.CDLL(name)
                break
            except OSError:
                continue
        if _rt is None:
            raise RuntimeError("HIP runtime library not found")
        _rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _rt.hipDeviceSynchronize.argtypes = []
        _rt.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _rt.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _rt.hipEventSynchronize.argtypes = [C.c_void_p]
        _rt.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _rt.hipEventDestroy.argtypes = [C.c_void_p]
    return _rt


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A float64 device buffer."""

    def __init__(self, shape):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        _ok(rt().hipMalloc(C.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        _ok(rt().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, 1), "hipMemcpy H2D")
        return d

    def to_host(self):
        a = np.empty(self.shape)
        _ok(rt().hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return a

    def free(self):
        if self.ptr is not None and self.ptr.value:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ok(rt().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.e = e

    def record(self, stream=None):
        _ok(rt().hipEventRecord(self.e, stream), "hipEventRecord")

    def elapsed_ms(self, later):
        _ok(rt().hipEventSynchronize(later.e), "hipEventSynchronize")
        t = C.c_float()
        _ok(rt().hipEventElapsedTime(C.byref(t), self.e, later.e), "hipEventElapsedTime")
        return float(t.value)


def copy_device(dst, src, nbytes):
    """Device-to-device copy (the runtime's own copy kernel)."""
    _ok(rt().hipMemcpy(dst.ptr, src.ptr, int(nbytes), 3), "hipMemcpy D2D")


def synchronize():
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
