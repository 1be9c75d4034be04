"""Minimal device-buffer handle over the C-ABI's allocator (no torch involved)."""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import check, lib


class DeviceArray:
    """A column-major device matrix (or flat buffer) of float64/float32."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().moao_hip_malloc(ctypes.byref(p), self.nbytes), "moao_hip_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().moao_hip_memcpy_h2d(ctypes.c_void_p(d.ptr), a.ctypes.data, d.nbytes),
              "moao_hip_memcpy_h2d")
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(lib().moao_hip_memcpy_d2h(out.ctypes.data, ctypes.c_void_p(self.ptr), self.nbytes),
              "moao_hip_memcpy_d2h")
        return out

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes <= self.nbytes
        check(lib().moao_hip_memcpy_d2d(ctypes.c_void_p(self.ptr), ctypes.c_void_p(other.ptr),
                                        other.nbytes), "moao_hip_memcpy_d2d")

    def free(self):
        if self.ptr:
            lib().moao_hip_free(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().moao_hip_synchronize(), "moao_hip_synchronize")


class Stream:
    """A library stream (default priority) for the stream-ordered entries."""

    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().moao_hip_stream_create(ctypes.byref(p)), "moao_hip_stream_create")
        self.ptr = p.value

    def synchronize(self):
        check(lib().moao_hip_stream_synchronize(ctypes.c_void_p(self.ptr)),
              "moao_hip_stream_synchronize")

    def destroy(self):
        if self.ptr:
            lib().moao_hip_stream_destroy(ctypes.c_void_p(self.ptr))
            self.ptr = None


class Profile:
    """Context manager collecting per-kernel-class event timings from the library."""

    def __enter__(self):
        check(lib().moao_hip_profile_begin(), "moao_hip_profile_begin")
        return self

    def __exit__(self, *exc):
        from .lib import PROF_CLASSES
        n = len(PROF_CLASSES)
        ms = (ctypes.c_double * n)()
        work = (ctypes.c_double * n)()
        cnt = (ctypes.c_long * n)()
        check(lib().moao_hip_profile_end(ms, work, cnt), "moao_hip_profile_end")
        self.result = {c: {"ms": ms[i], "work": work[i], "count": cnt[i]}
                       for i, c in enumerate(PROF_CLASSES)}
        return False
