"""Device memory, streams and events through the HIP runtime that
libhwtree.so itself links (ROCm's libamdhip64.so.7), via ctypes.

PyTorch is deliberately not used on the GPU path: its wheel bundles a second,
different HIP runtime, and two HIP runtimes in one process do not share
devices.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi

_hip = None


def hip():
    global _hip
    if _hip is None:
        capi.lib()  # loads libhwtree.so and, with it, its HIP runtime
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        L = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so.7"))
        P = C.c_void_p
        L.hipMalloc.argtypes = [C.POINTER(P), C.c_size_t]
        L.hipFree.argtypes = [P]
        L.hipMemcpy.argtypes = [P, P, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [P, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipStreamCreate.argtypes = [C.POINTER(P)]
        L.hipStreamDestroy.argtypes = [P]
        L.hipStreamSynchronize.argtypes = [P]
        L.hipEventCreate.argtypes = [C.POINTER(P)]
        L.hipEventDestroy.argtypes = [P]
        L.hipEventRecord.argtypes = [P, P]
        L.hipEventSynchronize.argtypes = [P]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), P, P]
        L.hipGetErrorString.argtypes = [C.c_int]
        L.hipGetErrorString.restype = C.c_char_p
        L.hipDeviceGetName = getattr(L, "hipDeviceGetName", None)
        _hip = L
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _ck(rc):
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}: {hip().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def set_device(i: int):
    _ck(hip().hipSetDevice(int(i)))


def synchronize():
    _ck(hip().hipDeviceSynchronize())


class DeviceArray:
    """A device buffer holding one numpy array's bytes."""

    def __init__(self, nbytes: int, dtype, shape):
        self.ptr = C.c_void_p()
        _ck(hip().hipMalloc(C.byref(self.ptr), max(int(nbytes), 1)))
        self.nbytes, self.dtype, self.shape = int(nbytes), np.dtype(dtype), tuple(shape)

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes, arr.dtype, arr.shape)
        if arr.nbytes:
            _ck(hip().hipMemcpy(d.ptr, arr.ctypes.data, arr.nbytes, H2D))
        return d

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        dt = np.dtype(dtype)
        return cls(int(np.prod(shape)) * dt.itemsize, dt, shape)

    @property
    def value(self) -> int:
        return self.ptr.value or 0

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _ck(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H))
        return out

    def fill_zero(self):
        _ck(hip().hipMemset(self.ptr, 0, self.nbytes))
        synchronize()  # the pricing plans run on their own non-blocking streams

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.e = C.c_void_p()
        _ck(hip().hipEventCreate(C.byref(self.e)))

    def record(self, stream=None):
        _ck(hip().hipEventRecord(self.e, stream))

    def synchronize(self):
        _ck(hip().hipEventSynchronize(self.e))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _ck(hip().hipEventElapsedTime(C.byref(ms), self.e, end.e))
        return float(ms.value)
