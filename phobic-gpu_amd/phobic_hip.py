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
    11: "PHOBIC_E_SCRATCH_LIMIT",
}
NOALLOC_SCRATCH_BYTES = 64 << 20
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
        ("alpha", ctypes.c_double),
    ]


class Outputs(ctypes.Structure):
    """phobic_outputs: host or device addresses of the build outputs."""
    _fields_ = [
        ("pilots", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("remap", ctypes.c_void_p),
        ("remap_offsets", ctypes.c_void_p),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("search_waves", ctypes.c_uint32),
        ("route_per", ctypes.c_uint32),
        ("max_fanout", ctypes.c_uint32),
        ("search_mode", ctypes.c_uint32),
        ("split_per", ctypes.c_uint32),
        ("first_fanout", ctypes.c_uint32),
        ("route_threads", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {"search_waves": self.search_waves, "route_per": self.route_per, "max_fanout": self.max_fanout,
                "search_mode": self.search_mode, "split_per": self.split_per, "first_fanout": self.first_fanout,
                "route_threads": self.route_threads}


class StageTimes(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("ms", ctypes.c_float * MAX_STAGES),
        ("names", (ctypes.c_char * 32) * MAX_STAGES),
        ("total_ms", ctypes.c_float),
        ("input_transfer_ms", ctypes.c_float),
        ("result_transfer_ms", ctypes.c_float),
    ]

    def as_dict(self) -> dict:
        d = {self.names[i].value.decode(): float(self.ms[i]) for i in range(self.count)}
        d["total"] = float(self.total_ms)
        if self.input_transfer_ms or self.result_transfer_ms:
            d["input_transfer"] = float(self.input_transfer_ms)
            d["result_transfer"] = float(self.result_transfer_ms)
        return d


# every symbol declared in include/phobic_hip.h
EXPORTS = [
    "phobic_ctx_create",
    "phobic_ctx_destroy",
    "phobic_status_string",
    "phobic_query_sizes",
    "phobic_remap_capacity",
    "phobic_fulcrums",
    "phobic_xxh3_128",
    "phobic_build_u64",
    "phobic_build_bytes",
    "phobic_build_keys",
    "phobic_build_u64_device",
    "phobic_build_bytes_device",
    "phobic_build_keys_device",
    "phobic_build_u64_device_noalloc",
    "phobic_batch_create",
    "phobic_batch_layout",
    "phobic_batch_launch",
    "phobic_batch_graph_launch",
    "phobic_batch_destroy",
    "phobic_server_create",
    "phobic_server_submit",
    "phobic_server_poll",
    "phobic_server_wait",
    "phobic_server_result",
    "phobic_server_fetch",
    "phobic_server_release",
    "phobic_server_stats_get",
    "phobic_server_destroy",
    "phobic_get_tuning",
    "phobic_set_tuning",
    "phobic_autotune",
    "phobic_last_stage_times",
    "phobic_eval_u64_device",
    "phobic_eval_bytes_device",
    "phobic_eval_pairs_device",
    "phobic_eval_keys_device",
    "phobic_check_bijection_device",
    "phobic_copy_device",
    "phobic_copy_to_device",
    "phobic_gen_splitmix64_device",
    "phobic_gen_strings_device",
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
    op = ctypes.POINTER(Outputs)
    sig = {
        "phobic_ctx_create": (I, [I, ctypes.POINTER(P)]),
        "phobic_ctx_destroy": (I, [P]),
        "phobic_status_string": (ctypes.c_char_p, [I]),
        "phobic_query_sizes": (I, [cfgp, U64, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
        "phobic_remap_capacity": (I, [cfgp, U64, ctypes.POINTER(U64)]),
        "phobic_fulcrums": (I, [cfgp, P]),
        "phobic_xxh3_128": (I, [P, U64, U64, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "phobic_build_u64": (I, [P, P, U64, cfgp, op, tp]),
        "phobic_build_bytes": (I, [P, P, P, U64, cfgp, op, tp]),
        "phobic_build_keys": (I, [P, P, U64, cfgp, op, tp]),
        "phobic_build_u64_device": (I, [P, P, U64, cfgp, op, P, P]),
        "phobic_build_bytes_device": (I, [P, P, P, U64, cfgp, op, P, P]),
        "phobic_build_keys_device": (I, [P, P, U64, cfgp, op, P, P]),
        "phobic_build_u64_device_noalloc": (I, [P, P, U64, cfgp, op, P, P]),
        "phobic_batch_create": (I, [P, cfgp, P, U32, ctypes.POINTER(P)]),
        "phobic_batch_layout": (I, [P, P, ctypes.POINTER(U32)]),
        "phobic_batch_launch": (I, [P, P, P, P, P, P]),
        "phobic_batch_graph_launch": (I, [P, P, P, P, P, P]),
        "phobic_batch_destroy": (I, [P]),
        "phobic_server_create": (I, [P, ctypes.POINTER(P)]),
        "phobic_server_submit": (I, [P, P, ctypes.POINTER(U64)]),
        "phobic_server_poll": (I, [P, U64, ctypes.POINTER(I)]),
        "phobic_server_wait": (I, [P, U64, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_double)]),
        "phobic_server_result": (I, [P, U64, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U64), op]),
        "phobic_server_fetch": (I, [P, U64, op]),
        "phobic_server_release": (I, [P, U64]),
        "phobic_server_stats_get": (I, [P, P]),
        "phobic_server_destroy": (I, [P]),
        "phobic_get_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "phobic_set_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "phobic_autotune": (I, [P, cfgp, U64, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_float)]),
        "phobic_last_stage_times": (I, [P, tp]),
        "phobic_eval_u64_device": (I, [P, P, U64, cfgp, U32, op, P, P]),
        "phobic_eval_bytes_device": (I, [P, P, P, U64, cfgp, U32, op, P, P]),
        "phobic_eval_pairs_device": (I, [P, P, U64, cfgp, U32, op, P, P]),
        "phobic_eval_keys_device": (I, [P, P, U64, cfgp, U32, op, P, P]),
        "phobic_check_bijection_device": (I, [P, P, U64, U64, P, P, P]),
        "phobic_copy_device": (I, [P, P, U64, P]),
        "phobic_copy_to_device": (I, [P, P, P, U64, P]),
        "phobic_gen_splitmix64_device": (I, [U64, U64, P, P]),
        "phobic_gen_strings_device": (I, [U64, U64, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def make_config(lam: float = 8.0, partition_size: int = 2048, seed: int = 0, alpha: float = 1.0) -> Config:
    """MPHFconfig(averageBucketSize, partitionSize) defaults (mphf_config.h:26)."""
    return Config(float(lam), int(partition_size), 0, int(seed), None, float(alpha))


def make_config_table(B: int, fulcrums: np.ndarray, lam: float = 8.0, partition_size: int = 2048, seed: int = 0,
                      alpha: float = 1.0) -> Config:
    """MPHFconfig(Bucketer*, ...) (mphf_config.h:34-42): a custom bucketer's
    B and 2048-entry fulcrum table.  The Config keeps the table alive."""
    f = np.ascontiguousarray(fulcrums, dtype=np.uint32)
    if f.size != 2048:
        raise ValueError("a fulcrum table has 2048 entries")
    cfg = Config(float(lam), int(partition_size), int(B), int(seed), f.ctypes.data, float(alpha))
    cfg._keep = f
    return cfg


def remap_capacity(cfg: Config, n: int) -> int:
    lib = load()
    v = ctypes.c_uint64()
    st = lib.phobic_remap_capacity(ctypes.byref(cfg), n, ctypes.byref(v))
    if st:
        raise PhobicError(st, "remap_capacity")
    return v.value


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
    remap: np.ndarray | None = None          # alpha < 1
    remap_offsets: np.ndarray | None = None  # alpha < 1


def device_outputs(pilots, offsets, remap=None, remap_offsets=None) -> Outputs:
    """phobic_outputs from device buffers (hip_rt.DeviceBuffer or addresses).
    The returned object keeps the buffer objects alive, so buffers created in
    the call are not freed while a build still writes to them."""
    out = Outputs(_ptr(pilots), _ptr(offsets), _ptr(remap) or None, _ptr(remap_offsets) or None)
    out._keep = (pilots, offsets, remap, remap_offsets)
    return out


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
        pilots = np.zeros(max(np_ * B, 1), dtype=np.uint32)
        offsets = np.zeros(np_ + 1, dtype=np.uint32)
        rcap = remap_capacity(cfg, n)
        remap = np.zeros(max(rcap, 1), dtype=np.uint32) if rcap else None
        roff = np.zeros(np_ + 1, dtype=np.uint32) if rcap else None
        out = Outputs(pilots.ctypes.data, offsets.ctypes.data, remap.ctypes.data if rcap else None,
                      roff.ctypes.data if rcap else None)
        t = StageTimes()
        st = fn(self.h, *args, ctypes.byref(cfg), ctypes.byref(out), ctypes.byref(t))
        if st:
            raise PhobicError(st, fn.__name__)
        r = BuildResult(pilots[:np_ * B], offsets, np_, B, t.as_dict())
        if rcap:
            r.remap, r.remap_offsets = remap[:int(roff[-1])], roff
        return r

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
    def build_u64_device(self, d_keys, n, cfg: Config, out: Outputs, d_status, stream=None) -> None:
        st = self.lib.phobic_build_u64_device(self.h, _ptr(d_keys), n, ctypes.byref(cfg), ctypes.byref(out),
                                              _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, "build_u64_device")

    def build_u64_device_noalloc(self, d_keys, n, cfg: Config, out: Outputs, d_status, stream=None) -> None:
        """phobic_build_u64_device_noalloc: the same build with the library's
        fixed per-device scratch array (no allocation in the call)."""
        st = self.lib.phobic_build_u64_device_noalloc(self.h, _ptr(d_keys), n, ctypes.byref(cfg), ctypes.byref(out),
                                                      _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, "build_u64_device_noalloc")

    def build_bytes_device(self, d_bytes, d_offs, n, cfg, out: Outputs, d_status, stream=None) -> None:
        st = self.lib.phobic_build_bytes_device(self.h, _ptr(d_bytes), _ptr(d_offs), n, ctypes.byref(cfg),
                                                ctypes.byref(out), _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, "build_bytes_device")

    def build_keys_device(self, d_keys, n, cfg, out: Outputs, d_status, stream=None) -> None:
        st = self.lib.phobic_build_keys_device(self.h, _ptr(d_keys), n, ctypes.byref(cfg), ctypes.byref(out),
                                               _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, "build_keys_device")

    def tuning(self) -> dict:
        t = Tuning()
        self.lib.phobic_get_tuning(self.h, ctypes.byref(t))
        return t.as_dict()

    def set_tuning(self, search_waves: int, route_per: int, max_fanout: int, search_mode: int = 0,
                   split_per: int = 0, first_fanout: int = 0, route_threads: int = 512) -> None:
        t = Tuning(search_waves, route_per, max_fanout, search_mode, split_per, first_fanout, route_threads)
        st = self.lib.phobic_set_tuning(self.h, ctypes.byref(t))
        if st:
            raise PhobicError(st, "set_tuning")

    def autotune(self, cfg: Config, sample_keys: int) -> tuple[dict, float]:
        t, ms = Tuning(), ctypes.c_float()
        st = self.lib.phobic_autotune(self.h, ctypes.byref(cfg), sample_keys, ctypes.byref(t), ctypes.byref(ms))
        if st:
            raise PhobicError(st, "autotune")
        return t.as_dict(), ms.value

    def last_stage_times(self) -> dict:
        t = StageTimes()
        st = self.lib.phobic_last_stage_times(self.h, ctypes.byref(t))
        if st:
            raise PhobicError(st, "last_stage_times")
        return t.as_dict()

    def copy_to_device(self, d_dst, host: np.ndarray, stream=None) -> None:
        """phobic_copy_to_device: a host array into device memory through the
        context's pinned staging chunks; returns when the copy is done."""
        host = np.ascontiguousarray(host)
        st = self.lib.phobic_copy_to_device(self.h, _ptr(d_dst), host.ctypes.data, host.nbytes, _stream(stream))
        if st:
            raise PhobicError(st, "copy_to_device")

    def eval_u64_device(self, d_keys, count, cfg, np_, mphf: Outputs, d_out, stream=None) -> None:
        st = self.lib.phobic_eval_u64_device(self.h, _ptr(d_keys), count, ctypes.byref(cfg), np_,
                                             ctypes.byref(mphf), _ptr(d_out), _stream(stream))
        if st:
            raise PhobicError(st, "eval_u64_device")

    def eval_bytes_device(self, d_bytes, d_offs, count, cfg, np_, mphf: Outputs, d_out, stream=None) -> None:
        st = self.lib.phobic_eval_bytes_device(self.h, _ptr(d_bytes), _ptr(d_offs), count, ctypes.byref(cfg), np_,
                                               ctypes.byref(mphf), _ptr(d_out), _stream(stream))
        if st:
            raise PhobicError(st, "eval_bytes_device")

    def eval_pairs_device(self, d_pairs, count, cfg, np_, mphf: Outputs, d_out, stream=None) -> None:
        st = self.lib.phobic_eval_pairs_device(self.h, _ptr(d_pairs), count, ctypes.byref(cfg), np_,
                                               ctypes.byref(mphf), _ptr(d_out), _stream(stream))
        if st:
            raise PhobicError(st, "eval_pairs_device")

    def eval_keys_device(self, d_keys, count, cfg, np_, mphf: Outputs, d_out, stream=None) -> None:
        st = self.lib.phobic_eval_keys_device(self.h, _ptr(d_keys), count, ctypes.byref(cfg), np_,
                                              ctypes.byref(mphf), _ptr(d_out), _stream(stream))
        if st:
            raise PhobicError(st, "eval_keys_device")

    def check_bijection_device(self, d_pos, count, n, d_bitmap, d_bad, stream=None) -> None:
        st = self.lib.phobic_check_bijection_device(
            self.h, _ptr(d_pos), count, n, _ptr(d_bitmap), _ptr(d_bad), _stream(stream))
        if st:
            raise PhobicError(st, "check_bijection_device")


KEYS_U64, KEYS_BYTES = 0, 1


class Request(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("keys_on_device", ctypes.c_int),
        ("keys", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("bytes_len", ctypes.c_uint64),
        ("n", ctypes.c_uint64),
        ("cfg", Config),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("device", ctypes.c_int),
        ("hbm_budget", ctypes.c_uint64),
        ("queue_capacity", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("keys_built", ctypes.c_uint64),
        ("hbm_in_use", ctypes.c_uint64),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("max_ms", ctypes.c_double),
        ("span_s", ctypes.c_double),
        ("grouped_launches", ctypes.c_uint64),
        ("grouped_requests", ctypes.c_uint64),
        ("pool_used", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Server:
    """phobic_server: continuous-batching builds with device-resident results.
    Host arrays passed to submit() are kept referenced until the request is
    released."""

    def __init__(self, slots: int = 4, hbm_budget: int = 200 << 30, device: int = 0, queue_capacity: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        cfg = ServerConfig(slots, device, hbm_budget, queue_capacity)
        st = self.lib.phobic_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if st:
            raise PhobicError(st, "server_create")
        self.h = h
        self._hold: dict[int, tuple] = {}
        self._cfgs: dict[int, Config] = {}

    def submit_u64(self, keys, cfg: Config, n: int | None = None) -> int:
        """keys: numpy u64 array (host) or a device buffer with n given."""
        on_dev = not isinstance(keys, np.ndarray)
        if not on_dev:
            keys = np.ascontiguousarray(keys, dtype=np.uint64)
            n = keys.size
        req = Request(KEYS_U64, int(on_dev), _ptr(keys) or None, None, 0, n, cfg)
        return self._submit(req, (keys,), cfg)

    def submit_bytes(self, data: np.ndarray, offsets: np.ndarray, cfg: Config) -> int:
        data = np.ascontiguousarray(data, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        buf = data if data.size else np.zeros(1, np.uint8)
        n = offsets.size - 1
        req = Request(KEYS_BYTES, 0, buf.ctypes.data, offsets.ctypes.data, int(data.size), n, cfg)
        return self._submit(req, (buf, offsets), cfg)

    def _submit(self, req: Request, hold: tuple, cfg: Config) -> int:
        rid = ctypes.c_uint64()
        st = self.lib.phobic_server_submit(self.h, ctypes.byref(req), ctypes.byref(rid))
        if st:
            raise PhobicError(st, "server_submit")
        self._hold[rid.value] = hold
        self._cfgs[rid.value] = cfg
        return rid.value

    def wait(self, rid: int) -> tuple[int, float]:
        st, lat = ctypes.c_int(), ctypes.c_double()
        e = self.lib.phobic_server_wait(self.h, rid, ctypes.byref(st), ctypes.byref(lat))
        if e:
            raise PhobicError(e, "server_wait")
        return st.value, lat.value

    def fetch(self, rid: int) -> BuildResult:
        np_, B, used = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        dev = Outputs()
        st = self.lib.phobic_server_result(self.h, rid, ctypes.byref(np_), ctypes.byref(B), ctypes.byref(used),
                                           ctypes.byref(dev))
        if st:
            raise PhobicError(st, "server_result")
        pil = np.zeros(max(np_.value * B.value, 1), np.uint32)
        off = np.zeros(np_.value + 1, np.uint32)
        has_remap = bool(dev.remap)
        rm = np.zeros(max(used.value, 1), np.uint32) if has_remap else None
        ro = np.zeros(np_.value + 1, np.uint32) if has_remap else None
        out = Outputs(pil.ctypes.data, off.ctypes.data, rm.ctypes.data if has_remap else None,
                      ro.ctypes.data if has_remap else None)
        st = self.lib.phobic_server_fetch(self.h, rid, ctypes.byref(out))
        if st:
            raise PhobicError(st, "server_fetch")
        r = BuildResult(pil[:np_.value * B.value], off, np_.value, B.value)
        if has_remap:
            r.remap, r.remap_offsets = rm[:used.value], ro
        return r

    def release(self, rid: int) -> None:
        st = self.lib.phobic_server_release(self.h, rid)
        if st:
            raise PhobicError(st, "server_release")
        self._hold.pop(rid, None)
        self._cfgs.pop(rid, None)

    def stats(self) -> dict:
        s = ServerStats()
        self.lib.phobic_server_stats_get(self.h, ctypes.byref(s))
        return s.as_dict()

    def close(self):
        if self.h:
            self.lib.phobic_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """phobic_batch: many independent u64 key sets built as one device build."""

    def __init__(self, ctx: "Context", cfg: Config, set_offsets: np.ndarray):
        self.lib = load()
        self.cfg = cfg
        self.set_offsets = np.ascontiguousarray(set_offsets, dtype=np.uint64)
        self.sets = self.set_offsets.size - 1
        h = ctypes.c_void_p()
        st = self.lib.phobic_batch_create(ctx.h, ctypes.byref(cfg), self.set_offsets.ctypes.data, self.sets,
                                          ctypes.byref(h))
        if st:
            raise PhobicError(st, "batch_create")
        self.h = h
        self.part_base = np.zeros(self.sets + 1, dtype=np.uint32)
        b = ctypes.c_uint32()
        self.lib.phobic_batch_layout(self.h, self.part_base.ctypes.data, ctypes.byref(b))
        self.B = b.value

    @property
    def total_pilots(self) -> int:
        return int(self.part_base[-1]) * self.B

    @property
    def total_offsets(self) -> int:
        return int(self.part_base[-1]) + self.sets

    def launch(self, d_keys, d_pilots, d_offsets, d_status, stream=None, graph: bool = False) -> None:
        fn = self.lib.phobic_batch_graph_launch if graph else self.lib.phobic_batch_launch
        st = fn(self.h, _ptr(d_keys), _ptr(d_pilots), _ptr(d_offsets), _ptr(d_status), _stream(stream))
        if st:
            raise PhobicError(st, fn.__name__)

    def split(self, pilots: np.ndarray, offsets: np.ndarray, s: int) -> tuple[np.ndarray, np.ndarray]:
        """Set s's pilots and offsets out of the batch outputs."""
        pb, pe = int(self.part_base[s]), int(self.part_base[s + 1])
        return pilots[self.B * pb:self.B * pe], offsets[pb + s:pe + s + 1]

    def close(self):
        if self.h:
            self.lib.phobic_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_device(d_dst, d_src, nbytes: int, stream=None) -> None:
    """phobic_copy_device: the library's plain streaming copy kernel."""
    lib = load()
    st = lib.phobic_copy_device(_ptr(d_dst), _ptr(d_src), nbytes, _stream(stream))
    if st:
        raise PhobicError(st, "copy_device")


def gen_splitmix64_device(seed: int, n: int, d_keys, stream=None) -> None:
    lib = load()
    st = lib.phobic_gen_splitmix64_device(seed, n, _ptr(d_keys), _stream(stream))
    if st:
        raise PhobicError(st, "gen_splitmix64_device")


def gen_strings_device(seed: int, n: int, d_offsets, d_bytes, stream=None) -> None:
    """Distinct printable strings in d_bytes at the ranges of d_offsets
    (n + 1 u64, every length >= 8); see phobic_gen_strings_device."""
    lib = load()
    st = lib.phobic_gen_strings_device(seed, n, _ptr(d_offsets), _ptr(d_bytes), _stream(stream))
    if st:
        raise PhobicError(st, "gen_strings_device")


def status_from_bits(bits: int) -> int:
    """Device status word (bit i set = error code i) -> lowest error code."""
    for c in range(1, 32):
        if bits & (1 << c):
            return c
    return PHOBIC_OK
