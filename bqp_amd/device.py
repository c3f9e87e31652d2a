"""Device buffers through the HIP runtime libbqp.so itself links (ctypes).

Keeps inputs resident in HBM for the device-pointer entry points without a
second HIP runtime in the process (torch wheels bundle their own
libamdhip64, which must not be mixed with ROCm's in one process).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _native.load()  # resolves libamdhip64.so.7 for libbqp.so first
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipSetDevice.argtypes = [C.c_int]
        _hip = L
    return _hip


def device_count() -> int:
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise _native.BqpError(f"hipDeviceSynchronize failed ({rc})")


class DeviceArray:
    """A device allocation holding a fixed-shape numpy dtype array."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._ptr = C.c_void_p()
        if self.nbytes:
            rc = hip().hipMalloc(C.byref(self._ptr), self.nbytes)
            if rc != 0:
                raise _native.BqpError(f"hipMalloc({self.nbytes}) failed ({rc})")

    @classmethod
    def from_numpy(cls, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.copy_from(a)
        return d

    @property
    def ptr(self) -> int:
        return self._ptr.value or 0

    def data_ptr(self) -> int:
        return self.ptr

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if self.nbytes and hip().hipMemcpy(self._ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, H2D) != 0:
            raise _native.BqpError("hipMemcpy H2D failed")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes and hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes, D2H) != 0:
            raise _native.BqpError("hipMemcpy D2H failed")
        return out

    def rows(self, start: int, count: int) -> "DeviceView":
        """Non-owning view of rows [start, start+count) along the first axis."""
        if start < 0 or count < 0 or start + count > self.shape[0]:
            raise ValueError("row range out of bounds")
        row_bytes = self.nbytes // self.shape[0] if self.shape[0] else 0
        return DeviceView(self.ptr + start * row_bytes, (count,) + self.shape[1:], self.dtype, self)

    def free(self):
        if self._ptr.value:
            hip().hipFree(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceView:
    """Rows of a DeviceArray (keeps the parent alive)."""

    def __init__(self, ptr: int, shape, dtype, parent):
        self._p, self.shape, self.dtype, self._parent = int(ptr), tuple(shape), np.dtype(dtype), parent
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize

    def data_ptr(self) -> int:
        return self._p

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes and hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self._p), self.nbytes,
                                           D2H) != 0:
            raise _native.BqpError("hipMemcpy D2H failed")
        return out
