"""Minimal HIP runtime access (device buffers, copies, synchronisation) for
driving the device entry points of libfmmh from Python without torch."""
import ctypes as ct
import os

import numpy as np

_hip = None


def runtime():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ct.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        _hip.hipFree.argtypes = [ct.c_void_p]
        _hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipStreamSynchronize.argtypes = [ct.c_void_p]
        _hip.hipGetDeviceCount.argtypes = [ct.POINTER(ct.c_int)]
        _hip.hipGetErrorString.restype = ct.c_char_p
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {runtime().hipGetErrorString(rc).decode()}")


def copy_d2d(dst_ptr, src_ptr, nbytes):
    """Device-to-device copy of nbytes between raw device addresses (ordered
    on the null stream; synchronize() before handing the data to a library
    stream)."""
    if nbytes:
        _check(runtime().hipMemcpy(dst_ptr, src_ptr, nbytes, D2D), "hipMemcpy D2D")


def device_count():
    n = ct.c_int(0)
    rc = runtime().hipGetDeviceCount(ct.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


def synchronize_copies():
    """Waits for the null stream only (the copies of copy_d2d), not for the
    library contexts' own (non-blocking) streams, whose work may still run."""
    _check(runtime().hipStreamSynchronize(None), "hipStreamSynchronize")


class DeviceArray:
    """A float64/uint32 device buffer with host copies."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ct.c_void_p()
        _check(runtime().hipMalloc(ct.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        _check(runtime().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other):
        """Device-to-device copy of other's bytes into this buffer."""
        n = min(self.nbytes, other.nbytes)
        _check(runtime().hipMemcpy(self.ptr, other.ptr, n, D2D), "hipMemcpy D2D")

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
