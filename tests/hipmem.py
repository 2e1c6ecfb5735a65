"""Device buffers for GPU tests through the HIP runtime's C API (ctypes).

The DEC entry points take caller device arrays; tests allocate them here,
copy host arrays in and out, and free them.  GPU tests only.
"""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        L.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        _hip = L
    return _hip


H2D, D2H = 1, 2


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """One FP64 device array with the shape of a host array."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.shape = host.shape
        self.nbytes = host.nbytes
        p = ctypes.c_void_p()
        _chk(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.shape == self.shape
        _chk(hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape)
        _chk(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
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


def ptrs(arrs):
    """A C array of device pointers (double *const *)."""
    return (ctypes.c_void_p * len(arrs))(*[a.ptr for a in arrs])


def synchronize():
    _chk(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A non-blocking HIP stream of default priority, created and destroyed here."""

    def __init__(self):
        p = ctypes.c_void_p()
        _chk(hip().hipStreamCreateWithFlags(ctypes.byref(p), 1), "hipStreamCreateWithFlags")
        self.handle = p.value

    def synchronize(self):
        _chk(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def copy_to_host(self, ptr, out):
        """Queue a device-to-host copy of out.nbytes bytes from ptr on this stream."""
        _chk(hip().hipMemcpyAsync(out.ctypes.data, ptr, out.nbytes, D2H, self.handle), "hipMemcpyAsync D2H")

    def destroy(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = None
