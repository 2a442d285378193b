"""Device buffers for callers of libfarm's device-pointer entry points.

The HIP runtime libfarm.so is linked against (libamdhip64) is loaded once
per process; this module binds its allocation and copy calls with ctypes so
that Python callers can hand device pointers to farm_intersect_dev,
farm_intersect_noalloc, farm_within_dev and farm_intersection_area_dev and
read the results back afterwards.  Nothing here computes anything.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi  # noqa: F401  (loads libfarm.so, and with it the HIP runtime)

_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.restype = C.c_int
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.restype = C.c_int
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.restype = C.c_int
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.restype = C.c_int
_hip.hipSetDevice.restype = C.c_int
_hip.hipSetDevice.argtypes = [C.c_int]

_H2D, _D2H = 1, 2


class DeviceBuffer:
    """nbytes of device memory on `device`, freed when the object goes."""

    def __init__(self, nbytes: int, device: int = 0):
        self.nbytes = max(int(nbytes), 1)
        self.device = device
        p = C.c_void_p()
        _check(_hip.hipSetDevice(device), "hipSetDevice")
        _check(_hip.hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self._p = p

    @property
    def ptr(self) -> int:
        return int(self._p.value)

    def to_host(self, dtype=np.uint8, count: int | None = None, offset: int = 0) -> np.ndarray:
        """count items of dtype starting `offset` bytes into the buffer."""
        dt = np.dtype(dtype)
        count = (self.nbytes - offset) // dt.itemsize if count is None else count
        out = np.empty(max(count, 1), dtype=dt)
        if count:
            _check(_hip.hipSetDevice(self.device), "hipSetDevice")
            _check(_hip.hipMemcpy(out.ctypes.data, self.ptr + offset, count * dt.itemsize, _D2H), "hipMemcpy")
        return out[:count]

    def from_host(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes:
            _check(_hip.hipSetDevice(self.device), "hipSetDevice")
            _check(_hip.hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy")

    def free(self) -> None:
        if getattr(self, "_p", None) is not None and self._p.value:
            _hip.hipFree(self._p)
        self._p = None

    def __del__(self):
        self.free()


def _check(rc: int, what: str) -> None:
    if rc:
        raise RuntimeError(f"{what} failed with HIP error {rc}")
