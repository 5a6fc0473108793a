"""Minimal HIP runtime access through ctypes (device memory, streams, events).

Loads the same libamdhip64 that libclkmeans.so links against, so device
pointers and streams created here are valid arguments to the library.
"""
import ctypes
import os

import numpy as np

_ROCM_LIB = os.environ.get("ROCM_PATH", "/opt/rocm") + "/lib/libamdhip64.so"
_hip = None

hipMemcpyHostToDevice = 1
hipMemcpyDeviceToHost = 2
hipMemcpyDeviceToDevice = 3


class HipError(RuntimeError):
    pass


def lib():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL(_ROCM_LIB, mode=ctypes.RTLD_GLOBAL)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [i]
        for name, args in {
            "hipGetDeviceCount": [ctypes.POINTER(i)],
            "hipSetDevice": [i],
            "hipMalloc": [ctypes.POINTER(vp), sz],
            "hipFree": [vp],
            "hipHostMalloc": [ctypes.POINTER(vp), sz, ctypes.c_uint],
            "hipHostFree": [vp],
            "hipMemcpy": [vp, vp, sz, i],
            "hipMemcpyAsync": [vp, vp, sz, i, vp],
            "hipMemsetAsync": [vp, i, sz, vp],
            "hipDeviceSynchronize": [],
            "hipStreamCreate": [ctypes.POINTER(vp)],
            "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp],
            "hipEventCreate": [ctypes.POINTER(vp)],
            "hipEventDestroy": [vp],
            "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
        }.items():
            fn = getattr(_hip, name)
            fn.argtypes = args
            fn.restype = i
    return _hip


def check(rc, what=""):
    if rc != 0:
        raise HipError("%s failed: %s" % (what, lib().hipGetErrorString(rc).decode()))


def device_count():
    n = ctypes.c_int(0)
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A hipMalloc'd buffer with a numpy dtype and shape."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_host(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.copy_from(arr, stream)
        return d

    def copy_from(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes, (arr.nbytes, self.nbytes)
        if stream is None:
            check(lib().hipMemcpy(self.ptr, arr.ctypes.data, self.nbytes, hipMemcpyHostToDevice), "hipMemcpy H2D")
        else:
            check(lib().hipMemcpyAsync(self.ptr, arr.ctypes.data, self.nbytes, hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync H2D")
            check(lib().hipStreamSynchronize(stream), "hipStreamSynchronize")

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, hipMemcpyDeviceToHost), "hipMemcpy D2H")
        return out

    @property
    def addr(self):
        return self.ptr.value

    def offset(self, elements):
        """Raw device address `elements` items past the start."""
        return ctypes.c_void_p(self.ptr.value + int(elements) * self.dtype.itemsize)

    def free(self):
        if self.ptr and self.ptr.value:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A numpy array in page-locked host memory (hipHostMalloc), so copies
    from it run asynchronously at full link speed."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        nbytes = max(int(np.prod(shape, dtype=np.int64)) * dtype.itemsize, 1)
        self.ptr = ctypes.c_void_p()
        check(lib().hipHostMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes), 0), "hipHostMalloc")
        buf = (ctypes.c_char * nbytes).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape, dtype=np.int64))).reshape(shape)

    def free(self):
        if self.ptr and self.ptr.value:
            self.array = None
            lib().hipHostFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                lib().hipStreamDestroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        check(lib().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().hipEventDestroy(self.handle)
        except Exception:
            pass
