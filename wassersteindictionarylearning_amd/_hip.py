"""Minimal ctypes access to the HIP runtime libwdl.so is linked against
(device buffers, copies and events for callers that hold no framework).

Loaded by soname after libwdl.so, so it is the same runtime instance.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from . import lib

_rt = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _rt
    if _rt is None:
        lib()  # pulls in the runtime libwdl.so was linked against
        r = ctypes.CDLL("libamdhip64.so.7")
        r.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        r.hipFree.argtypes = [ctypes.c_void_p]
        r.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        r.hipDeviceSynchronize.argtypes = []
        r.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        r.hipEventDestroy.argtypes = [ctypes.c_void_p]
        r.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        r.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        r.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                          ctypes.c_void_p]
        r.hipGetErrorString.restype = ctypes.c_char_p
        r.hipGetErrorString.argtypes = [ctypes.c_int]
        _rt = r
    return _rt


def _ok(e, what):
    if e != 0:
        raise RuntimeError("%s: %s" % (what, rt().hipGetErrorString(e).decode()))


class DeviceArray:
    """A device allocation holding a copy of a numpy array (or empty)."""

    def __init__(self, shape, dtype=np.float32, host=None):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _ok(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    @classmethod
    def from_host(cls, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype=dtype)
        return cls(a.shape, dtype, a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _ok(rt().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _ok(rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.free()
            except Exception:
                pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _ok(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.h = e.value

    def record(self, stream: int):
        _ok(rt().hipEventRecord(self.h, stream), "hipEventRecord")

    def synchronize(self):
        _ok(rt().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _ok(rt().hipEventElapsedTime(ctypes.byref(ms), self.h, end.h), "hipEventElapsedTime")
        return float(ms.value)

    def destroy(self):
        if self.h:
            rt().hipEventDestroy(self.h)
            self.h = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.destroy()
            except Exception:
                pass


def synchronize():
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
