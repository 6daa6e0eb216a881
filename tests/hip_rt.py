"""Minimal device-memory helpers for tests, bound from the HIP runtime that libcugip_hip.so links
(libamdhip64.so.7 from /opt/rocm). torch ships a second HIP runtime; two runtimes in one process
do not share devices, so the tests keep torch out of processes that load the library."""
import ctypes

import cugip_amd  # noqa: F401  (loads libamdhip64.so.7 first)

_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, host):
        import numpy as np
        self.host = np.ascontiguousarray(host)
        self.ptr = ctypes.c_void_p()
        if _hip.hipMalloc(ctypes.byref(self.ptr), self.host.nbytes) != 0:
            raise MemoryError("hipMalloc failed")
        if _hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy failed")

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()
