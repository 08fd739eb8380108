"""Device memory of the HIP path without a framework runtime: DeviceArray, device_count,
synchronize. (PyTorch bundles its own HIP runtime; the product uses ROCm's through
libdyng_hip.so only, so the two never share a process.)"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def device_count() -> int:
    c = C.c_int(0)
    _lib.check(_lib.lib().dyng_hip_device_count(C.byref(c)))
    return c.value


def synchronize(device: int = 0) -> None:
    _lib.check(_lib.lib().dyng_hip_synchronize(device))


class DeviceArray:
    """A contiguous array in device memory (hipMalloc), filled from a host array."""

    __slots__ = ("ptr", "shape", "dtype", "device")

    def __init__(self, host, dtype=None, device: int = 0):
        a = np.ascontiguousarray(np.asarray(host, dtype=dtype))
        self.shape = a.shape
        self.dtype = a.dtype
        self.device = device
        p = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_malloc(device, a.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        if a.nbytes:
            _lib.check(_lib.lib().dyng_hip_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    @classmethod
    def empty(cls, shape, dtype, device: int = 0) -> "DeviceArray":
        """An uninitialised device array (no host copy is made)."""
        self = object.__new__(cls)
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.device = device
        self.ptr = 0
        p = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_malloc(device, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        return self

    @property
    def size(self) -> int:
        return int(np.prod(self.shape)) if self.shape else 1

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            _lib.check(_lib.lib().dyng_hip_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def __del__(self):
        p = getattr(self, "ptr", 0)
        if p and _lib._lib is not None:
            _lib._lib.dyng_hip_free(p)
            self.ptr = 0

    def __repr__(self) -> str:
        return f"dyng_amd.DeviceArray(shape={self.shape}, dtype={self.dtype})"
