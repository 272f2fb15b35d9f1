"""Minimal HIP runtime access (device buffers, streams, events) through the
same libamdhip64 instance that libptamd.so is linked against, so pointers and
stream handles passed to the C-ABI come from one runtime.  Plumbing for tests
and benchmarks; importing it does not touch the GPU."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as _ptlib  # noqa: F401  (loads libptamd first: fixes which libamdhip64 binds)

_hip = ctypes.CDLL("libamdhip64.so.7")
_vp = ctypes.c_void_p
_H2D, _D2H, _D2D = 1, 2, 3

for _n, _a in {
    "hipMalloc": [ctypes.POINTER(_vp), ctypes.c_size_t],
    "hipFree": [_vp],
    "hipMemcpy": [_vp, _vp, ctypes.c_size_t, ctypes.c_int],
    "hipMemset": [_vp, ctypes.c_int, ctypes.c_size_t],
    "hipDeviceSynchronize": [],
    "hipSetDevice": [ctypes.c_int],
    "hipGetDeviceCount": [ctypes.POINTER(ctypes.c_int)],
    "hipStreamCreate": [ctypes.POINTER(_vp)],
    "hipStreamDestroy": [_vp],
    "hipStreamSynchronize": [_vp],
    "hipEventCreate": [ctypes.POINTER(_vp)],
    "hipEventDestroy": [_vp],
    "hipEventRecord": [_vp, _vp],
    "hipEventSynchronize": [_vp],
    "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), _vp, _vp],
    "hipGetErrorString": [ctypes.c_int],
}.items():
    _f = getattr(_hip, _n)
    _f.argtypes = _a
    _f.restype = ctypes.c_int
_hip.hipGetErrorString.restype = ctypes.c_char_p


class HipError(RuntimeError):
    pass


def _ok(rc, what):
    if rc != 0:
        raise HipError(f"{what}: {_hip.hipGetErrorString(rc).decode()}")


def device_count():
    c = ctypes.c_int(0)
    if _hip.hipGetDeviceCount(ctypes.byref(c)) != 0:
        return 0
    return c.value


def synchronize():
    _ok(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A hipMalloc'd buffer with a numpy shape and dtype."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _ok(_hip.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(_hip.hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        return d

    def to_host(self):
        a = np.empty(self.shape, self.dtype)
        _ok(_hip.hipMemcpy(a.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy D2H")
        return a

    def copy_from(self, other):
        """Device-to-device copy of min(nbytes) bytes (the runtime's blit kernel)."""
        _ok(_hip.hipMemcpy(self.ptr, other.ptr, min(self.nbytes, other.nbytes), _D2D), "hipMemcpy D2D")

    def zero(self):
        _ok(_hip.hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def fill_bytes(self, value):
        """Every byte set to `value` (0..255)."""
        _ok(_hip.hipMemset(self.ptr, int(value), self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = _vp()
        _ok(_hip.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            _hip.hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = _vp()
        _ok(_hip.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        _ok(_hip.hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        _ok(_hip.hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        _ok(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle:
            _hip.hipEventDestroy(self.handle)
            self.handle = None
