"""Minimal HIP runtime access for device buffers, streams and events.

Binds libamdhip64.so.7 by SONAME, i.e. the same runtime instance liblbx.so
uses in this process (whichever copy the dynamic linker loaded first), so
pointers, streams and events are interchangeable with the C ABI.
"""
import ctypes

import numpy as np

_hip = None


def runtime():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, i32]
        h.hipMemset.argtypes = [vp, i32, sz]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        h.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipEventDestroy.argtypes = [vp]
        h.hipEventRecord.argtypes = [vp, vp]
        h.hipEventSynchronize.argtypes = [vp]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        h.hipGetErrorString.argtypes = [i32]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _ok(err, what):
    if err != 0:
        raise RuntimeError("%s failed: %s" % (what, runtime().hipGetErrorString(err).decode()))


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """A hipMalloc'd byte range; data_ptr() makes it usable as PolynomialMatrix storage."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _ok(runtime().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 16)), "hipMalloc")

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr)
        return b

    def data_ptr(self):
        return self.ptr.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _ok(runtime().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype=np.uint32, count=None):
        count = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        _ok(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        _ok(runtime().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    def __init__(self):
        self.ev = ctypes.c_void_p()
        _ok(runtime().hipEventCreate(ctypes.byref(self.ev)), "hipEventCreate")

    def record(self, stream=None):
        _ok(runtime().hipEventRecord(self.ev, stream), "hipEventRecord")

    def synchronize(self):
        _ok(runtime().hipEventSynchronize(self.ev), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _ok(runtime().hipEventElapsedTime(ctypes.byref(ms), self.ev, end.ev), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            runtime().hipEventDestroy(self.ev)
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(runtime().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _ok(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    @property
    def value(self):
        return self.handle.value

    def destroy(self):
        """Destroy the stream after waiting for its work.  A blocking stream
        that outlives its use costs every later launch on the null stream."""
        if self.handle.value:
            runtime().hipStreamSynchronize(self.handle)
            runtime().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
