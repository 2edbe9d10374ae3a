"""Minimal HIP runtime plumbing (device buffers, copies, synchronisation) through ctypes.

Used to keep benchmark inputs resident in HBM before the timed region; the E-step itself runs in
librln_estep.so.
"""
import ctypes as C

import numpy as np

_hip = C.CDLL("libamdhip64.so")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipDeviceSynchronize.argtypes = []
_hip.hipSetDevice.argtypes = [C.c_int]
_hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
_hip.hipGetErrorString.argtypes = [C.c_int]
_hip.hipGetErrorString.restype = C.c_char_p
_hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]

H2D, D2H, D2D = 1, 2, 3


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {_hip.hipGetErrorString(rc).decode()}")


def device_count():
    n = C.c_int(0)
    rc = _hip.hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def set_device(d):
    _check(_hip.hipSetDevice(d), "hipSetDevice")


def synchronize():
    _check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    _check(_hip.hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


def copy_device(dst, src, nbytes):
    """Device-to-device copy (runs the runtime's copy kernel): used to calibrate traffic counters."""
    _check(_hip.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, D2D), "hipMemcpy D2D")


class DeviceArray:
    """A device copy of a numpy array (freed with the object)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.shape, self.dtype = host.shape, host.dtype
        self.ptr = C.c_void_p()
        _check(_hip.hipMalloc(C.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")
        _check(_hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    @property
    def address(self):
        return self.ptr.value

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _check(_hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
