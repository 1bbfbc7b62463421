"""Minimal ctypes binding of the HIP runtime (libamdhip64) for device buffers,
streams and events.

The library and this module load the same libamdhip64.so from /opt/rocm, so a
pointer from DeviceArray can be passed straight to the C ABI of libabhsf.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
_hip = None

H2D, D2H, D2D = 1, 2, 3


def _load():
    global _hip
    if _hip is None:
        path = os.path.join(_ROCM, "lib", "libamdhip64.so")
        _hip = C.CDLL(path, mode=C.RTLD_GLOBAL)
        vp, sz, i, f = C.c_void_p, C.c_size_t, C.c_int, C.c_float
        _hip.hipMalloc.argtypes = [C.POINTER(vp), sz]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, sz, i]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, sz, i, vp]
        _hip.hipMemset.argtypes = [vp, i, sz]
        _hip.hipMemsetAsync.argtypes = [vp, i, sz, vp]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(i)]
        _hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipEventCreate.argtypes = [C.POINTER(vp)]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(f), vp, vp]
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipGetErrorString.argtypes = [i]
        _hip.hipDeviceGetName.argtypes = [C.c_char_p, i, i]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(sz), C.POINTER(sz)]
    return _hip


class HipError(RuntimeError):
    pass


def check(rc: int, what: str = "hip call") -> None:
    if rc != 0:
        msg = _load().hipGetErrorString(rc).decode()
        raise HipError(f"{what} failed: {msg} ({rc})")


def device_count() -> int:
    """Number of visible GPUs (0 when the runtime finds none)."""
    try:
        h = _load()
    except OSError:
        return 0
    n = C.c_int(0)
    if h.hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def device_name(dev: int = 0) -> str:
    buf = C.create_string_buffer(256)
    check(_load().hipDeviceGetName(buf, 256, dev), "hipDeviceGetName")
    return buf.value.decode()


def synchronize() -> None:
    check(_load().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = C.c_size_t(0), C.c_size_t(0)
    check(_load().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """Owning device buffer with a numpy dtype/shape."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = C.c_void_p(0)
        if self.nbytes:
            check(_load().hipMalloc(C.byref(self.ptr), self.nbytes), f"hipMalloc({self.nbytes})")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(_load().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, H2D), "hipMemcpy H2D")
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float64) -> "DeviceArray":
        d = cls(shape, dtype)
        if d.nbytes:
            check(_load().hipMemset(d.ptr, 0, d.nbytes), "hipMemset")
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            check(_load().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(_load().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def fill_zero(self, stream=None) -> None:
        if self.nbytes:
            check(_load().hipMemsetAsync(self.ptr, 0, self.nbytes, stream), "hipMemsetAsync")

    def copy_from(self, other: "DeviceArray", stream=None) -> None:
        assert other.nbytes == self.nbytes
        if self.nbytes:
            check(_load().hipMemcpyAsync(self.ptr, other.ptr, self.nbytes, D2D, stream), "hipMemcpyAsync D2D")

    @property
    def address(self) -> int:
        return self.ptr.value or 0

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            _load().hipFree(self.ptr)
            self.ptr = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        self.handle = C.c_void_p(0)
        check(_load().hipStreamCreate(C.byref(self.handle)), "hipStreamCreate")

    def synchronize(self) -> None:
        check(_load().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                _load().hipStreamDestroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = C.c_void_p(0)
        check(_load().hipEventCreate(C.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None) -> None:
        s = stream.handle if isinstance(stream, Stream) else stream
        check(_load().hipEventRecord(self.handle, s), "hipEventRecord")

    def synchronize(self) -> None:
        check(_load().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0.0)
        check(_load().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _load().hipEventDestroy(self.handle)
        except Exception:
            pass
