"""Device buffers for GPU tests that hand device pointers to the C-ABI,
through the HIP runtime libsdpd.so itself links (torch bundles a HIP
runtime of its own, and two runtimes cannot share a process)."""
import ctypes as C

import numpy as np

from mcf_amd.sdpd import load_library

_hip = None


def hip():
    global _hip
    if _hip is None:
        load_library()  # maps libamdhip64.so.7, which the name below then finds
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = h
    return _hip


_H2D, _D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (or room for one of a shape/dtype)."""

    def __init__(self, host=None, shape=None, dtype=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            shape, dtype = host.shape, host.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        if host is not None:
            self.put(host)

    def put(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy to the device failed ({rc})")

    def get(self):
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy from the device failed ({rc})")
        return out

    @property
    def addr(self) -> int:
        return int(self.ptr.value)

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")
