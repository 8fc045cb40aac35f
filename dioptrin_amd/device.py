"""Device arrays owned through the C ABI (dio_device_alloc / dio_memcpy).

The library runs on the HIP runtime it was built against; callers that have
no device allocator of their own (tests, benchmarks, the serving loop) keep
frames and results in these buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class DeviceArray:
    """A contiguous device buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype, device: int = 0):
        from .integrator import _Context
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.device = device
        self._ctx = _Context.get(device)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        N.check(N.lib().dio_device_alloc(self._ctx, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        self._owner = None

    @classmethod
    def from_numpy(cls, a: np.ndarray, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        if d.nbytes:
            N.check(N.lib().dio_memcpy(d._ctx, d.ptr, a.ctypes.data, d.nbytes, 0))
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            N.check(N.lib().dio_memcpy(self._ctx, out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def __getitem__(self, k: slice) -> "DeviceArray":
        """A view of frames [start, stop) along the first axis."""
        if not isinstance(k, slice) or k.step not in (None, 1):
            raise TypeError("only contiguous slices of the first axis are supported")
        start, stop, _ = k.indices(self.shape[0])
        view = DeviceArray.__new__(DeviceArray)
        view.shape = (max(0, stop - start),) + self.shape[1:]
        view.dtype, view.device, view._ctx = self.dtype, self.device, self._ctx
        row = int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize
        view.nbytes = view.shape[0] * row
        view.ptr = self.ptr + start * row
        view._owner = self
        return view

    def zero_(self):
        self.copy_from(np.zeros(self.shape, self.dtype))

    def copy_from(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        N.check(N.lib().dio_memcpy(self._ctx, self.ptr, a.ctypes.data, self.nbytes, 0))

    def __del__(self):
        if getattr(self, "_owner", 1) is None and getattr(self, "ptr", 0):
            N.lib().dio_device_free(self._ctx, self.ptr)
            self.ptr = 0


class PinnedArray:
    """A numpy array in page-locked host memory (the pinned staging of the
    reference's GPU path, cuda/device.rs:1063-1079): dio_integrate copies
    such frames to the device by DMA. `.array` is the numpy view; it is
    valid while this object lives."""

    def __init__(self, shape, dtype, device: int = 0):
        from .integrator import _Context
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self._ctx = _Context.get(device)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        N.check(N.lib().dio_host_alloc(self._ctx, self.nbytes, C.byref(p)))
        self.ptr = p.value or 0
        buf = (C.c_char * self.nbytes).from_address(self.ptr) if self.nbytes else bytearray()
        self.array = np.frombuffer(buf, dtype=self.dtype).reshape(self.shape)

    def __del__(self):
        if getattr(self, "ptr", 0):
            self.array = None
            N.lib().dio_host_free(self._ctx, self.ptr)
            self.ptr = 0


def synchronize(device: int = 0):
    from .integrator import _Context
    N.check(N.lib().dio_synchronize(_Context.get(device)))


def synthetic_poisson_u32(rate: np.ndarray, backgrounds, seed: int, device: int = 0) -> DeviceArray:
    """Frames Poisson(rate + background[f]) generated on the device."""
    r = DeviceArray.from_numpy(np.ascontiguousarray(rate, np.float32), device)
    bg = np.ascontiguousarray(backgrounds, np.float64)
    out = DeviceArray((bg.size,) + tuple(rate.shape), np.uint32, device)
    N.check(N.lib().dio_synthetic_poisson_u32(r._ctx, r.ptr, rate.size, bg.ctypes.data, bg.size,
                                              seed, out.ptr))
    return out
