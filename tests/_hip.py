"""Device buffers and streams for the GPU tests through the HIP runtime libceos_amd itself
links (/opt/rocm/lib/libamdhip64.so.7, found by its soname), so the tests need no second
runtime in the process (torch ships its own)."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from ceos_amd import capi

        capi.lib()  # loads libceos_amd.so and, with it, the runtime it links
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of a shape and dtype)."""

    def __init__(self, a=None, shape=None, dtype=None):
        if a is not None:
            a = np.ascontiguousarray(a)
            shape, dtype = a.shape, a.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(1, self.nbytes)), "hipMalloc")
        self.ptr = p.value
        if a is not None:
            _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def at(self, row):
        """Device address of row `row` (first axis)."""
        return self.ptr + row * (self.nbytes // self.shape[0])

    def fill(self, byte):
        _ok(hip().hipMemset(self.ptr, byte, self.nbytes), "hipMemset")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        if getattr(self, "handle", None):
            hip().hipStreamDestroy(self.handle)
            self.handle = None


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
