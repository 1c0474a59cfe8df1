"""Device buffers for GPU tests through libamdhip64 (no framework needed)."""
import ctypes

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
    return _hip


class DeviceBuffer:
    def __init__(self, host_bytes):
        self.ptr = ctypes.c_void_p()
        n = len(host_bytes)
        assert hip().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(n)) == 0
        assert hip().hipMemcpy(self.ptr, host_bytes.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n), 1) == 0

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None
