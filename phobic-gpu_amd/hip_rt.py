"""Minimal ctypes access to the HIP runtime that libphobic_hip.so links
(/opt/rocm libamdhip64.so.7): device buffers, streams and events for tests
and benchmarks.

PyTorch ships a second, separately versioned HIP runtime; loading both into
one process makes whichever initialises second fail, so the test and bench
processes use this module instead of torch for device memory.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_rt = None

H2D, D2H, D2D = 1, 2, 3


def rt() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        # the same soname the library resolves, so both share one runtime
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7"):
            try:
                _rt = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _rt is None:
            raise RuntimeError("HIP runtime libamdhip64.so.7 not found")
        P, S, I = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        for name, args in {
            "hipMalloc": [ctypes.POINTER(P), S],
            "hipFree": [P],
            "hipMemcpy": [P, P, S, I],
            "hipMemcpyAsync": [P, P, S, I, P],
            "hipMemsetAsync": [P, I, S, P],
            "hipStreamCreate": [ctypes.POINTER(P)],
            "hipStreamSynchronize": [P],
            "hipStreamDestroy": [P],
            "hipDeviceSynchronize": [],
            "hipEventCreate": [ctypes.POINTER(P)],
            "hipEventRecord": [P, P],
            "hipEventSynchronize": [P],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), P, P],
            "hipEventDestroy": [P],
            "hipGetDeviceCount": [ctypes.POINTER(I)],
            "hipSetDevice": [I],
            "hipDeviceGetName": [ctypes.c_char_p, I, I],
            "hipMemGetInfo": [ctypes.POINTER(S), ctypes.POINTER(S)],
        }.items():
            fn = getattr(_rt, name)
            fn.restype = I
            fn.argtypes = args
    return _rt


def check(err: int, what: str) -> None:
    if err != 0:
        raise RuntimeError(f"{what} failed with hipError {err}")


def device_count() -> int:
    n = ctypes.c_int(0)
    if rt().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def device_name(dev: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    check(rt().hipDeviceGetName(buf, 256, dev), "hipDeviceGetName")
    return buf.value.decode()


def mem_info() -> tuple[int, int]:
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    check(rt().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


def synchronize() -> None:
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd byte buffer."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), f"hipMalloc({nbytes})")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            check(rt().hipMemcpy(b.ptr, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")
        return b

    def to_numpy(self, dtype, count: int | None = None) -> np.ndarray:
        dt = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dt.itemsize
        out = np.empty(count, dtype=dt)
        if out.nbytes:
            check(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self, stream=None) -> None:
        check(rt().hipMemsetAsync(self.ptr, 0, self.nbytes, stream), "hipMemsetAsync")

    def free(self) -> None:
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(rt().hipStreamCreate(ctypes.byref(p)), "hipStreamCreate")
        self.handle = p.value

    def synchronize(self) -> None:
        check(rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __int__(self) -> int:
        return self.handle or 0

    def destroy(self) -> None:
        if self.handle:
            rt().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(p)), "hipEventCreate")
        self.handle = p.value

    def record(self, stream=None) -> None:
        check(rt().hipEventRecord(self.handle, int(stream) if stream is not None else None), "hipEventRecord")

    def synchronize(self) -> None:
        check(rt().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(rt().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value
