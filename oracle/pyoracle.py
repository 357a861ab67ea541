"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg; never by the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

ORC_OK = 0


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
    lib = ctypes.CDLL(LIB_PATH)
    P, U32, U64, D = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double
    sig = {
        "orc_xxh3_128": (None, [P, ctypes.c_size_t, U64, P, P]),
        "orc_hash_u64": (None, [P, U64, U64, P]),
        "orc_hash_bytes": (None, [P, P, U64, U64, P]),
        "orc_bucket_count": (U32, [D, U32]),
        "orc_fulcrums": (None, [D, U32, P]),
        "orc_hi32_mul": (U32, [U32, U32]),
        "orc_partition_of": (U32, [U32, U32]),
        "orc_bucket_of": (U32, [U32, P]),
        "orc_mix32": (U32, [U32, U32]),
        "orc_build_keys": (ctypes.c_int, [P, U64, D, U32, P, P, P, P]),
        "orc_build_keys_ex": (ctypes.c_int, [P, U64, D, U32, U32, P, P, P, P, P, P]),
        "orc_query_keys_ex": (None, [P, U64, P, P, U32, U32, P, U32, P, P, P]),
        "orc_table_size": (U32, [U32, U32]),
        "orc_remap_capacity": (U64, [U64, U32, U32]),
        "orc_sizes": (None, [U64, D, U32, P, P]),
        "orc_query_keys": (None, [P, U64, P, P, U32, U32, P, P]),
        "orc_compact_words": (U64, [U64, U64]),
        "orc_compact_width": (U64, [P, U64]),
        "orc_compact_encode": (U64, [P, U64, P, P]),
        "orc_compact_access": (U64, [P, U64, U64]),
        "orc_encode_mphf": (U64, [P, P, U32, U32, U32, P, U64, U64, U32, P, P]),
        "orc_rice_param": (U64, [P, U64]),
        "orc_build_keys_table": (ctypes.c_int, [P, U64, U32, U32, P, U32, P, P, P, P]),
        "orc_scan_splitmix": (ctypes.c_int, [U64, U64, U64, D, U32, P, P]),
        "orc_build_splitmix_parts": (ctypes.c_int, [U64, U64, U64, D, U32, U32, P, U32, P, P, P, U64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start+1 .. start+n of splitmix64 from state `seed`
    (state_i = seed + i * 0x9E3779B97F4A7C15, then the 30/27/31 mix)."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def xxh3_128(data: bytes, seed: int = 0) -> int:
    lib = load()
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    lib.orc_xxh3_128(data, len(data), seed, ctypes.byref(lo), ctypes.byref(hi))
    return (hi.value << 64) | lo.value


def hash_u64(keys: np.ndarray, seed: int = 0) -> np.ndarray:
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty((keys.size, 4), dtype=np.uint32)
    lib.orc_hash_u64(keys.ctypes.data, keys.size, seed, out.ctypes.data)
    return out


def hash_bytes(data: np.ndarray, offsets: np.ndarray, seed: int = 0) -> np.ndarray:
    lib = load()
    data = np.ascontiguousarray(data, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = offsets.size - 1
    out = np.empty((n, 4), dtype=np.uint32)
    buf = data if data.size else np.zeros(1, np.uint8)
    lib.orc_hash_bytes(buf.ctypes.data, offsets.ctypes.data, n, seed, out.ctypes.data)
    return out


def sizes(n: int, lam: float, P: int) -> tuple[int, int]:
    lib = load()
    a, b = ctypes.c_uint32(), ctypes.c_uint32()
    lib.orc_sizes(n, lam, P, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def fulcrums(lam: float, P: int) -> np.ndarray:
    lib = load()
    f = np.zeros(2048, dtype=np.uint32)
    lib.orc_fulcrums(lam, P, f.ctypes.data)
    return f


def build_keys(keys: np.ndarray, lam: float, P: int, intermediates: bool = False):
    """Returns (status, pilots, offsets[, bucket_sizes, order])."""
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 4)
    n = keys.shape[0]
    np_, B = sizes(n, lam, P)
    pilots = np.zeros(np_ * B, dtype=np.uint32)
    offsets = np.zeros(np_ + 1, dtype=np.uint32)
    bs = np.zeros(np_ * B, dtype=np.uint32) if intermediates else None
    order = np.zeros(np_ * B, dtype=np.uint32) if intermediates else None
    kp = keys.ctypes.data if n else np.zeros(4, np.uint32).ctypes.data
    st = lib.orc_build_keys(kp, n, lam, P, pilots.ctypes.data, offsets.ctypes.data,
                            bs.ctypes.data if bs is not None else None,
                            order.ctypes.data if order is not None else None)
    if intermediates:
        return st, pilots, offsets, bs, order
    return st, pilots, offsets


def alpha_ppm(alpha: float) -> int:
    return int(round(alpha * 1_000_000))


def build_keys_alpha(keys: np.ndarray, lam: float, P: int, alpha: float):
    """alpha <= 1 build: (status, pilots, offsets, remap, remap_offsets)."""
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 4)
    n = keys.shape[0]
    np_, B = sizes(n, lam, P)
    a = alpha_ppm(alpha)
    pilots = np.zeros(np_ * B, dtype=np.uint32)
    offsets = np.zeros(np_ + 1, dtype=np.uint32)
    remap = np.zeros(max(1, lib.orc_remap_capacity(n, np_, a)), dtype=np.uint32)
    roff = np.zeros(np_ + 1, dtype=np.uint32)
    kp = keys.ctypes.data if n else np.zeros(4, np.uint32).ctypes.data
    st = lib.orc_build_keys_ex(kp, n, lam, P, a, pilots.ctypes.data, offsets.ctypes.data, remap.ctypes.data,
                               roff.ctypes.data, None, None)
    used = int(roff[-1]) if a < 1_000_000 else 0
    return st, pilots, offsets, remap[:used], roff


def bucketer_fulcrums(kind: str, lam: float, P: int) -> tuple[int, np.ndarray]:
    """(B, table) of MPHFconfig(Bucketer*, lambda, P) for the Uniform or Skew
    bucketer: B = round(P / lambda), f[i] = (u32)(y(i/2047) * (B << 16)),
    f[0] = 0, f[2047] = B << 16 (mphf_config.h:34-55, bucketers.h:18-50)."""
    B = int(np.floor(P / lam + 0.5))  # C round(): halves away from zero
    y = {"uniform": lambda x: x,
         "skew": lambda x: x / 0.6 * 0.3 if x < 0.6 else 0.3 + (x - 0.6) / 0.4 * 0.7}[kind]
    f = np.zeros(2048, np.uint32)
    for i in range(1, 2047):
        f[i] = int(y(i / 2047.0) * float(B << 16))
    f[2047] = B << 16
    return B, f


def build_keys_table(keys: np.ndarray, P: int, B: int, fulcs: np.ndarray, alpha: float = 1.0):
    """Build with a custom fulcrum table: (status, pilots, offsets, remap, remap_offsets)."""
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 4)
    n = keys.shape[0]
    np_ = (n + P - 1) // P
    a = alpha_ppm(alpha)
    pilots = np.zeros(np_ * B, dtype=np.uint32)
    offsets = np.zeros(np_ + 1, dtype=np.uint32)
    remap = np.zeros(max(1, lib.orc_remap_capacity(n, np_, a)), dtype=np.uint32)
    roff = np.zeros(np_ + 1, dtype=np.uint32)
    f = np.ascontiguousarray(fulcs, dtype=np.uint32)
    kp = keys.ctypes.data if n else np.zeros(4, np.uint32).ctypes.data
    st = lib.orc_build_keys_table(kp, n, P, B, f.ctypes.data, a, pilots.ctypes.data, offsets.ctypes.data,
                                  remap.ctypes.data, roff.ctypes.data)
    used = int(roff[-1]) if a < 1_000_000 else 0
    return st, pilots, offsets, remap[:used], roff


def query_keys_alpha(keys, pilots, offsets, remap, roff, lam, P, alpha):
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 4)
    np_ = offsets.size - 1
    B = pilots.size // max(np_, 1)
    f = fulcrums(lam, P)
    out = np.zeros(keys.shape[0], dtype=np.uint32)
    rm = np.ascontiguousarray(remap, dtype=np.uint32) if remap.size else np.zeros(1, np.uint32)
    lib.orc_query_keys_ex(keys.ctypes.data, keys.shape[0], np.ascontiguousarray(pilots).ctypes.data,
                          np.ascontiguousarray(offsets).ctypes.data, np_, B, f.ctypes.data, alpha_ppm(alpha),
                          rm.ctypes.data, np.ascontiguousarray(roff).ctypes.data, out.ctypes.data)
    return out


def query_keys(keys: np.ndarray, pilots: np.ndarray, offsets: np.ndarray, lam: float, P: int) -> np.ndarray:
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.uint32).reshape(-1, 4)
    np_ = offsets.size - 1
    B = pilots.size // max(np_, 1)
    f = fulcrums(lam, P)
    out = np.zeros(keys.shape[0], dtype=np.uint32)
    lib.orc_query_keys(keys.ctypes.data, keys.shape[0], np.ascontiguousarray(pilots).ctypes.data,
                       np.ascontiguousarray(offsets).ctypes.data, np_, B, f.ctypes.data, out.ctypes.data)
    return out


def scan_splitmix(kseed: int, n: int, hseed: int, lam: float, P: int, max_bucket: bool = True):
    """Per-partition key counts and largest bucket of the keys splitmix64(kseed)
    1..n, hashed with XXH3 seed hseed, without materialising them."""
    lib = load()
    np_, _ = sizes(n, lam, P)
    counts = np.zeros(max(np_, 1), np.uint32)
    mb = np.zeros(max(np_, 1), np.uint16) if max_bucket else None
    st = lib.orc_scan_splitmix(kseed, n, hseed, lam, P, counts.ctypes.data, mb.ctypes.data if mb is not None else None)
    if st:
        raise RuntimeError(f"orc_scan_splitmix: status {st}")
    return counts[:np_], (mb[:np_] if mb is not None else None)


def build_splitmix_parts(kseed: int, n: int, hseed: int, lam: float, P: int, alpha: float, parts):
    """Builds only the listed partitions of the splitmix64(kseed) key stream:
    (status, pilots[len(parts), B], counts[len(parts)], remap entries
    concatenated in list order)."""
    lib = load()
    parts = np.ascontiguousarray(parts, dtype=np.uint32)
    np_, B = sizes(n, lam, P)
    a = alpha_ppm(alpha)
    k = parts.size
    pil = np.zeros((max(k, 1), B), np.uint32)
    cnt = np.zeros(max(k, 1), np.uint32)
    cap = int(lib.orc_remap_capacity(k * (P + P // 2), k, a)) if a < 1_000_000 else 0
    rm = np.zeros(max(cap, 1), np.uint32)
    st = lib.orc_build_splitmix_parts(kseed, n, hseed, lam, P, a, parts.ctypes.data if k else None, k,
                                      pil.ctypes.data, cnt.ctypes.data, rm.ctypes.data, cap)
    used = int(sum(int(lib.orc_table_size(int(c), a)) - int(c) for c in cnt[:k])) if (a < 1_000_000 and st == 0) else 0
    return st, pil[:k], cnt[:k], rm[:used]


def compact_encode(values: np.ndarray) -> tuple[np.ndarray, int]:
    lib = load()
    values = np.ascontiguousarray(values, dtype=np.uint64)
    w = lib.orc_compact_width(values.ctypes.data, values.size)
    words = np.zeros(lib.orc_compact_words(values.size, w), dtype=np.uint64)
    wo = ctypes.c_uint64()
    lib.orc_compact_encode(values.ctypes.data, values.size, words.ctypes.data, ctypes.byref(wo))
    return words, wo.value


def compact_access(words: np.ndarray, width: int, i: int) -> int:
    return load().orc_compact_access(np.ascontiguousarray(words).ctypes.data, width, i)


# pilot layouts: interleaved columns (FastQueryMphf compact, SmallSpaceMphf
# rice), one sequence (mono_encoders.hpp), dual columns (the reference
# benchmark's default interleaved_encoder_dual<rice, compact>, split 0.5)
ENCODINGS = {"compact": 0, "rice": 1, "mono_compact": 2, "mono_rice": 3, "dual_rice_compact": 6}
# partition offsets: differences to i*P (diff_partition_offset_encoder.hpp) or
# the offsets themselves (direct_partition_offset_encoder.hpp)
OFFSET_ENCODINGS = {"diff_compact": 0, "diff_rice": 1, "direct_compact": 2, "direct_rice": 3}


def encode_mphf(pilots: np.ndarray, offsets: np.ndarray, lam: float, P: int, n: int, seed: int = 0,
                encoding: str = "compact", with_bits_per_key: bool = False,
                offset_encoding: str = "diff_compact"):
    """Serialised MPHF bytes ("PHOBMPH2"); with_bits_per_key also returns the
    bits-per-key figure."""
    lib = load()
    np_ = offsets.size - 1
    B = pilots.size // max(np_, 1) if np_ else sizes(n, lam, P)[1]
    f = fulcrums(lam, P)
    pil = np.ascontiguousarray(pilots, dtype=np.uint32)
    off = np.ascontiguousarray(offsets, dtype=np.uint32)
    pp = pil.ctypes.data if pil.size else np.zeros(1, np.uint32).ctypes.data
    enc = ENCODINGS[encoding] | (OFFSET_ENCODINGS[offset_encoding] << 8)
    size = lib.orc_encode_mphf(pp, off.ctypes.data, np_, B, P, f.ctypes.data, n, seed, enc, None, None)
    out = np.zeros(size, dtype=np.uint8)
    bpk = ctypes.c_double()
    lib.orc_encode_mphf(pp, off.ctypes.data, np_, B, P, f.ctypes.data, n, seed, enc, out.ctypes.data,
                        ctypes.byref(bpk))
    return (out.tobytes(), bpk.value) if with_bits_per_key else out.tobytes()


def rice_param(values: np.ndarray) -> int:
    v = np.ascontiguousarray(values, dtype=np.uint64)
    return int(load().orc_rice_param(v.ctypes.data if v.size else np.zeros(1, np.uint64).ctypes.data, v.size))
