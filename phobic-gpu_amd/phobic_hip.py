"""ctypes binding of libphobic_hip.so (the C ABI in include/phobic_hip.h).

Thin plumbing for tests and benchmarks: numpy arrays for the host-buffer
entry points, hip_rt.DeviceBuffer objects (or raw device addresses) and
hip_rt.Stream handles for the device entry points.  Loading fails loudly
when the HIP library has not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphobic_hip.so")

PHOBIC_OK = 0
STATUS_NAMES = {
    0: "PHOBIC_OK",
    1: "PHOBIC_E_INVALID_CONFIG",
    2: "PHOBIC_E_HIP",
    3: "PHOBIC_E_OOM",
    4: "PHOBIC_E_NO_DEVICE",
    5: "PHOBIC_E_PARTITION_OVERFLOW",
    6: "PHOBIC_E_BUCKET_OVERFLOW",
    7: "PHOBIC_E_DUPLICATE_KEYS",
    8: "PHOBIC_E_PILOT_OVERFLOW",
    9: "PHOBIC_E_BUSY",
    10: "PHOBIC_E_NOT_FOUND",
}
MAX_STAGES = 8


class PhobicError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")


class Config(ctypes.Structure):
    _fields_ = [
        ("lambda_", ctypes.c_double),
        ("partition_size", ctypes.c_uint32),
        ("buckets", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("fulcrums", ctypes.c_void_p),
    ]


class StageTimes(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("ms", ctypes.c_float * MAX_STAGES),
        ("names", (ctypes.c_char * 32) * MAX_STAGES),
        ("total_ms", ctypes.c_float),
    ]

    def as_dict(self) -> dict:
        d = {self.names[i].value.decode(): float(self.ms[i]) for i in range(self.count)}
        d["total"] = float(self.total_ms)
        return d


# every symbol declared in include/phobic_hip.h
EXPORTS = [
    "phobic_ctx_create",
    "phobic_ctx_destroy",
    "phobic_status_string",
    "phobic_query_sizes",
    "phobic_fulcrums",
    "phobic_xxh3_128",
    "phobic_build_u64",
    "phobic_build_bytes",
    "phobic_build_keys",
    "phobic_build_u64_device",
    "phobic_build_bytes_device",
    "phobic_build_keys_device",
    "phobic_last_stage_times",
    "phobic_eval_u64_device",
    "phobic_check_bijection_device",
    "phobic_gen_splitmix64_device",
]

_lib = None


def load() -> ctypes.CDLL:
    """Load libphobic_hip.so and declare argument types; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, U32, U64, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    cfgp = ctypes.POINTER(Config)
    tp = ctypes.POINTER(StageTimes)
    sig = {
        "phobic_ctx_create": (I, [I, ctypes.POINTER(P)]),
        "phobic_ctx_destroy": (I, [P]),
        "phobic_status_string": (ctypes.c_char_p, [I]),
        "phobic_query_sizes": (I, [cfgp, U64, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
        "phobic_fulcrums": (I, [cfgp, P]),
        "phobic_xxh3_128": (I, [P, U64, U64, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "phobic_build_u64": (I, [P, P, U64, cfgp, P, P, tp]),
        "phobic_build_bytes": (I, [P, P, P, U64, cfgp, P, P, tp]),
        "phobic_build_keys": (I, [P, P, U64, cfgp, P, P, tp]),
        "phobic_build_u64_device": (I, [P, P, U64, cfgp, P, P, P, P]),
        "phobic_build_bytes_device": (I, [P, P, P, U64, cfgp, P, P, P, P]),
        "phobic_build_keys_device": (I, [P, P, U64, cfgp, P, P, P, P]),
        "phobic_last_stage_times": (I, [P, tp]),
        "phobic_eval_u64_device": (I, [P, P, U64, cfgp, U32, P, P, P, P]),
        "phobic_check_bijection_device": (I, [P, P, U64, U64, P, P, P]),
        "phobic_gen_splitmix64_device": (I, [U64, U64, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def make_config(lam: float = 8.0, partition_size: int = 2048, seed: int = 0) -> Config:
    """MPHFconfig(averageBucketSize, partitionSize) defaults (mphf_config.h:26)."""
    return Config(float(lam), int(partition_size), 0, int(seed), None)


def sizes(cfg: Config, n: int) -> tuple[int, int]:
    lib = load()
    np_, b = ctypes.c_uint32(), ctypes.c_uint32()
    st = lib.phobic_query_sizes(ctypes.byref(cfg), n, ctypes.byref(np_), ctypes.byref(b))
    if st:
        raise PhobicError(st, "query_sizes")
    return np_.value, b.value


def fulcrums(cfg: Config) -> np.ndarray:
    lib = load()
    out = np.zeros(2048, dtype=np.uint32)
    st = lib.phobic_fulcrums(ctypes.byref(cfg), out.ctypes.data)
    if st:
        raise PhobicError(st, "fulcrums")
    return out


def _ptr(a) -> int:
    """numpy array (host), hip_rt.DeviceBuffer, or a raw integer address."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, int):
        return a
    return int(a.ptr)


def _stream(stream=None) -> int:
    return int(stream) if stream is not None else 0


@dataclass
class BuildResult:
    pilots: np.ndarray
    offsets: np.ndarray
    np_: int
    B: int
    times: dict = field(default_factory=dict)


class Context:
    """One phobic_ctx (device memory pool + stream)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        st = self.lib.phobic_ctx_create(device, ctypes.byref(h))
        if st:
            raise PhobicError(st, "ctx_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.phobic_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- host-buffer builds ----
    def _host(self, fn, args, cfg: Config, n: int) -> BuildResult:
        np_, B = sizes(cfg, n)
        pilots = np.zeros(np_ * B, dtype=np.uint32)
        offsets = np.zeros(np_ + 1, dtype=np.uint32)
        t = StageTimes()
        st = fn(self.h, *args, ctypes.byref(cfg), pilots.ctypes.data, offsets.ctypes.data, ctypes.byref(t))
        if st:
            raise PhobicError(st, fn.__name__)
        return BuildResult(pilots, offsets, np_, B, t.as_dict())

    def build_u64(self, keys: np.ndarray, cfg: Config) -> BuildResult:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        return self._host(self.lib.phobic_build_u64, (keys.ctypes.data, keys.size), cfg, keys.size)

    def build_keys(self, keys: np.ndarray, cfg: Config) -> BuildResult:
        keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 4)
        return self._host(self.lib.phobic_build_keys, (keys.ctypes.data, keys.shape[0]), cfg, keys.shape[0])

    def build_bytes(self, data: np.ndarray, offsets: np.ndarray, cfg: Config) -> BuildResult:
        data = np.ascontiguousarray(data, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = offsets.size - 1
        buf = data if data.size else np.zeros(1, dtype=np.uint8)
        return self._host(self.lib.phobic_build_bytes, (buf.ctypes.data, offsets.ctypes.data, n), cfg, n)

    # ---- device builds (device buffers; enqueue on `stream`, default null) ----
    def build_u64_device(self, d_keys, n, cfg: Config, d_pilots, d_offsets, d_status, stream=None) -> None:
        st = self.lib.phobic_build_u64_device(
            self.h, _ptr(d_keys), n, ctypes.byref(cfg), _ptr(d_pilots), _ptr(d_offsets),
            _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, "build_u64_device")

    def build_bytes_device(self, d_bytes, d_offs, n, cfg, d_pilots, d_offsets, d_status, stream=None) -> None:
        st = self.lib.phobic_build_bytes_device(
            self.h, _ptr(d_bytes), _ptr(d_offs), n, ctypes.byref(cfg), _ptr(d_pilots), _ptr(d_offsets),
            _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, "build_bytes_device")

    def build_keys_device(self, d_keys, n, cfg, d_pilots, d_offsets, d_status, stream=None) -> None:
        st = self.lib.phobic_build_keys_device(
            self.h, _ptr(d_keys), n, ctypes.byref(cfg), _ptr(d_pilots), _ptr(d_offsets), _ptr(d_status),
            _stream(stream))
        if st:
            raise PhobicError(st, "build_keys_device")

    def last_stage_times(self) -> dict:
        t = StageTimes()
        st = self.lib.phobic_last_stage_times(self.h, ctypes.byref(t))
        if st:
            raise PhobicError(st, "last_stage_times")
        return t.as_dict()

    def eval_u64_device(self, d_keys, count, cfg, np_, d_pilots, d_offsets, d_out, stream=None) -> None:
        st = self.lib.phobic_eval_u64_device(
            self.h, _ptr(d_keys), count, ctypes.byref(cfg), np_, _ptr(d_pilots), _ptr(d_offsets),
            _ptr(d_out), _stream(stream))
        if st:
            raise PhobicError(st, "eval_u64_device")

    def check_bijection_device(self, d_pos, count, n, d_bitmap, d_bad, stream=None) -> None:
        st = self.lib.phobic_check_bijection_device(
            self.h, _ptr(d_pos), count, n, _ptr(d_bitmap), _ptr(d_bad), _stream(stream))
        if st:
            raise PhobicError(st, "check_bijection_device")


def gen_splitmix64_device(seed: int, n: int, d_keys, stream=None) -> None:
    lib = load()
    st = lib.phobic_gen_splitmix64_device(seed, n, _ptr(d_keys), _stream(stream))
    if st:
        raise PhobicError(st, "gen_splitmix64_device")


def status_from_bits(bits: int) -> int:
    """Device status word (bit i set = error code i) -> lowest error code."""
    for c in range(1, 32):
        if bits & (1 << c):
            return c
    return PHOBIC_OK
