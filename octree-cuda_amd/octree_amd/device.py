"""Device memory through the HIP runtime (libamdhip64), for callers that hand the library clouds
already resident in HBM (ocpi_set_point_cloud_device). Plain allocation and copies; no framework."""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None

H2D, D2H = 1, 2


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        lib = ctypes.CDLL("libamdhip64.so")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _hip = lib
    return _hip


class DeviceArray:
    """nbytes of device memory; freed by close()."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        rc = hip().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(1, self.nbytes)))
        if rc != 0:
            raise MemoryError(f"hipMalloc of {nbytes} bytes failed ({rc})")

    def address(self, offset: int = 0) -> int:
        return self.ptr.value + offset

    def upload(self, host: np.ndarray, offset: int = 0):
        host = np.ascontiguousarray(host)
        if offset + host.nbytes > self.nbytes:
            raise ValueError("upload beyond the device array")
        rc = hip().hipMemcpy(ctypes.c_void_p(self.address(offset)), host.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(host.nbytes), H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy to the device failed ({rc})")

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
