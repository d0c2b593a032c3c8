"""Device arrays for the _device entry points of the library (dj_ics_lpt_device, dj_deposit_device,
dj_density_loss_device, ...): memory allocated with the HIP runtime the library itself links, so that
arrays can stay on the GPU between calls (the white noise and its cotangent of a gradient loop, a loss
target).  Julia callers pass ROCArray pointers instead."""
import ctypes as C

import numpy as np

from . import _lib

_H2D, _D2H, _D2D = 1, 2, 3


class DeviceArray:
    """A contiguous array in device memory: shape, dtype, device, nbytes, ptr (the device address)."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.device = int(device)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._h = _lib.hip()
        self._p = C.c_void_p()
        if self._h.hipSetDevice(self.device) != 0 or self._h.hipMalloc(C.byref(self._p), max(self.nbytes, 1)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes on device {self.device} failed")

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.copy_from(a)
        return d

    @property
    def ptr(self):
        return self._p.value

    def copy_from(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"{a.nbytes} bytes into a device array of {self.nbytes}")
        self._h.hipSetDevice(self.device)
        if self._h.hipMemcpy(self._p, a.ctypes.data, self.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        self._h.hipSetDevice(self.device)
        if self._h.hipMemcpy(out.ctypes.data, self._p, self.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def free(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            self._h.hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.free()
