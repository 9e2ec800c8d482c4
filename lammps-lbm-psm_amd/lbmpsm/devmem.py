"""Device buffers from the HIP runtime that liblbmpsm.so itself links.

PyTorch wheels bundle their own libamdhip64; a process that initialises
that runtime first and then the library's leaves the second without a
device.  Buffers handed to lbmpsm_set_particles_device are therefore
allocated here, through the library's own runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        N.LIB  # loads liblbmpsm.so and, with it, its libamdhip64
        lib = C.CDLL("libamdhip64.so.7")
        lib.hipMalloc.restype = C.c_int
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipFree.restype = C.c_int
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipMemcpy.restype = C.c_int
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipDeviceSynchronize.restype = C.c_int
        lib.hipSetDevice.restype = C.c_int
        lib.hipSetDevice.argtypes = [C.c_int]
        _HIP = lib
    return _HIP


def device_count() -> int:
    """HIP devices visible to the library's runtime (0 when none)."""
    hip = _hip()
    hip.hipGetDeviceCount.restype = C.c_int
    hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
    n = C.c_int(0)
    return n.value if hip.hipGetDeviceCount(C.byref(n)) == 0 else 0


class DeviceArray:
    """A copy of a numpy array in device memory (freed on close / GC)."""

    def __init__(self, host: np.ndarray, device: int = 0):
        hip = _hip()
        a = np.ascontiguousarray(host)
        self.dtype = a.dtype
        self.shape = a.shape
        self.nbytes = a.nbytes
        self.ptr = C.c_void_p()
        if hip.hipSetDevice(device) != 0:
            raise RuntimeError("hipSetDevice failed")
        if hip.hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed")
        if self.nbytes and hip.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy host->device failed")

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def to_host(self) -> np.ndarray:
        """A host copy (the device is synchronised first)."""
        hip = _hip()
        out = np.empty(self.shape, dtype=self.dtype)
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
        if self.nbytes and hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy device->host failed")
        return out

    def close(self):
        if self.ptr and self.ptr.value:
            _hip().hipDeviceSynchronize()
            _hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
