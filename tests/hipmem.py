"""Device buffers for tests that call the C ABI's device-pointer entry points: hipMalloc /
hipMemcpy of the HIP runtime libgpvmc_hip.so itself links (one runtime per process)."""
import ctypes

import numpy as np

from gpvmc_amd import binding

_HIP = None
H2D, D2H = 1, 2


def hip():
    global _HIP
    if _HIP is None:
        binding.load_library()  # loads the runtime the library was linked against
        h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipSetDevice.argtypes = [ctypes.c_int]
        _HIP = h
    return _HIP


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    """A device allocation holding a copy of a numpy array's bytes (or zeros)."""

    def __init__(self, like, dtype=None, shape=None):
        if isinstance(like, np.ndarray):
            self.host = np.ascontiguousarray(like)
        else:
            self.host = np.zeros(shape, dtype)
        self.nbytes = max(self.host.nbytes, 1)
        p = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value
        if isinstance(like, np.ndarray):
            check(hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D), "hipMemcpy")

    def get(self):
        out = np.empty_like(self.host)
        check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
