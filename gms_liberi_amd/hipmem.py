"""Device buffers, streams and events through the HIP runtime the library links.

Loading libliberi_hip.so pulls in the system HIP runtime (libamdhip64);
these helpers call that same runtime through ctypes, so device-resident
builds need no second GPU stack in the process.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .liberi import load_library

_hip = None


def runtime() -> C.CDLL:
    global _hip
    if _hip is None:
        load_library()  # resolves libamdhip64 first, from the library's own link
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventQuery.argtypes = [C.c_void_p]
        h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        h.hipHostFree.argtypes = [C.c_void_p]
        h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        h.hipGetErrorString.restype = C.c_char_p
        _hip = h
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


class DeviceArray:
    """A float64 array in HBM."""

    def __init__(self, n: int):
        self.n = n
        self.ptr = C.c_void_p()
        _ok(runtime().hipMalloc(C.byref(self.ptr), max(1, n) * 8), "hipMalloc")

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, np.float64)
        d = cls(a.shape[0])
        _ok(runtime().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n)
        _ok(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self) -> None:
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


    def copy_from_async(self, host: "PinnedArray", stream: "Stream") -> None:
        _ok(runtime().hipMemcpyAsync(self.ptr, host.ptr, self.n * 8, 1, stream.ptr), "hipMemcpyAsync H2D")

    def copy_to_async(self, host: "PinnedArray", stream: "Stream") -> None:
        _ok(runtime().hipMemcpyAsync(host.ptr, self.ptr, self.n * 8, 2, stream.ptr), "hipMemcpyAsync D2H")


class PinnedArray:
    """A float64 array in page-locked host memory (hipHostMalloc), for
    asynchronous copies; .array is a numpy view of it."""

    def __init__(self, n: int):
        self.n = n
        self.ptr = C.c_void_p()
        _ok(runtime().hipHostMalloc(C.byref(self.ptr), max(1, n) * 8, 0), "hipHostMalloc")
        self.array = np.ctypeslib.as_array(C.cast(self.ptr, C.POINTER(C.c_double)), shape=(n,))

    def free(self) -> None:
        if self.ptr:
            self.array = None
            runtime().hipHostFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_memory() -> tuple[int, int]:
    """(free, total) bytes of HBM on the current device."""
    free, total = C.c_size_t(), C.c_size_t()
    _ok(runtime().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class Stream:
    def __init__(self):
        self.ptr = C.c_void_p()
        _ok(runtime().hipStreamCreate(C.byref(self.ptr)), "hipStreamCreate")

    @property
    def address(self) -> int:
        return self.ptr.value

    def synchronize(self) -> None:
        _ok(runtime().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.ptr:
            runtime().hipStreamDestroy(self.ptr)
            self.ptr = C.c_void_p()


class Event:
    def __init__(self):
        self.ptr = C.c_void_p()
        _ok(runtime().hipEventCreate(C.byref(self.ptr)), "hipEventCreate")

    def record(self, stream: Stream) -> None:
        _ok(runtime().hipEventRecord(self.ptr, stream.ptr), "hipEventRecord")

    def synchronize(self) -> None:
        _ok(runtime().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def done(self) -> bool:
        rc = runtime().hipEventQuery(self.ptr)
        if rc == 600:  # hipErrorNotReady
            return False
        _ok(rc, "hipEventQuery")
        return True

    def destroy(self) -> None:
        if self.ptr:
            runtime().hipEventDestroy(self.ptr)
            self.ptr = C.c_void_p()

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _ok(runtime().hipEventElapsedTime(C.byref(ms), self.ptr, end.ptr), "hipEventElapsedTime")
        return ms.value
