"""Writes the golden fixtures under tests/golden/.

Run from the repository root in a container where the pins are available:
    python tests/golden/make_golden.py

* xxh3_vectors.json -- XXH3-128 digests from the python `xxhash` package
  (3.8.1, bundling xxHash 0.8.2).  The reference calls XXH128(ptr, len, 0)
  (include/phobicGpu/hasher.hpp:34,43,52,61) from its xxHash submodule, which
  is not vendored; XXH3-128 output is frozen across the 0.8.x line.  Inputs
  are regenerated by `vector_bytes(length, data_seed)` below, so only the
  digests are stored.
* fulcrums.npz -- MPHFconfig fulcrum tables and bucket counts printed by
  oracle/_ref/ref_config, which is compiled from the reference's own
  include/phobicGpu/mphf_config.h + bucketers.h by `make -C oracle ref`.
* c1_pilots.npz -- the oracle's pilots/offsets for the 1M-key configuration
  (splitmix64 seed 1, lambda 7.0, P 2048, hash seed 0); a regression pin for
  the oracle itself, which tests also check for bijection.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLD = os.path.join(ROOT, "tests", "golden")

LENGTHS = list(range(0, 257)) + [300, 511, 512, 513, 1023, 1024, 1025, 2047, 2048, 2049, 4096, 5000]
SEEDS = [0, 42, 0xFFFFFFFFFFFFFFFF, 0x9E3779B97F4A7C15]

CONFIGS = [(7.0, 2048), (8.0, 2500), (7.5, 2048), (6.0, 2048), (4.5, 2500), (8.0, 2048), (7.0, 100), (3.0, 64)]


def vector_bytes(length: int, data_seed: int) -> bytes:
    """Deterministic input bytes: splitmix64 stream from data_seed, little-endian."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    words = pyoracle.splitmix64(data_seed, (length + 7) // 8)
    return words.astype("<u8").tobytes()[:length]


def make_xxh3() -> None:
    import xxhash

    vecs = []
    for i, ln in enumerate(LENGTHS):
        for j, seed in enumerate(SEEDS):
            ds = 1000 * i + j
            data = vector_bytes(ln, ds)
            vecs.append({"len": ln, "data_seed": ds, "seed": str(seed),
                         "digest": "%032x" % xxhash.xxh3_128_intdigest(data, seed)})
    # u64 keys hashed as their 8 native bytes (hasher.hpp:42-48)
    keys = []
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    for k in [int(x) for x in pyoracle.splitmix64(1, 64)] + [0, 1, 2**64 - 1]:
        for seed in (0, 42):
            keys.append({"key": str(k), "seed": seed,
                         "digest": "%032x" % xxhash.xxh3_128_intdigest(k.to_bytes(8, "little"), seed)})
    with open(os.path.join(GOLD, "xxh3_vectors.json"), "w") as f:
        json.dump({"source": f"python xxhash {xxhash.VERSION} (xxHash {xxhash.XXHASH_VERSION})",
                   "bytes": vecs, "u64": keys}, f, indent=0)


def make_fulcrums() -> None:
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_config")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {}
    for lam, P in CONFIGS:
        lines = subprocess.run([exe, repr(lam), str(P)], check=True, capture_output=True, text=True).stdout.split("\n")
        B, maxsize = map(int, lines[0].split())
        out[f"f_{lam}_{P}"] = np.array(list(map(int, lines[1].split())), dtype=np.uint32)
        out[f"b_{lam}_{P}"] = np.array([B, maxsize], dtype=np.uint32)
    # MPHFconfig(Bucketer*, lambda, P) with the uniform and skew bucketers
    for kind in ("uniform", "skew"):
        for lam, P in ((7.0, 2048), (4.5, 1000)):
            lines = subprocess.run([exe, repr(lam), str(P), kind], check=True, capture_output=True,
                                   text=True).stdout.split("\n")
            out[f"{kind}_{lam}_{P}"] = np.array(list(map(int, lines[1].split())), dtype=np.uint32)
            out[f"{kind}b_{lam}_{P}"] = np.array([int(lines[0].split()[0])], dtype=np.uint32)
    np.savez_compressed(os.path.join(GOLD, "fulcrums.npz"), **out)


def make_c1() -> None:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    keys = pyoracle.splitmix64(1, 1_000_000)
    hk = pyoracle.hash_u64(keys, 0)
    st, pil, off = pyoracle.build_keys(hk, 7.0, 2048)
    assert st == 0
    np.savez_compressed(os.path.join(GOLD, "c1_pilots.npz"), pilots=pil, offsets=off)


if __name__ == "__main__":
    which = sys.argv[1:] or ["xxh3", "fulcrums", "c1"]
    if "xxh3" in which:
        make_xxh3()
    if "fulcrums" in which:
        make_fulcrums()
    if "c1" in which:
        make_c1()
    print("golden fixtures written to", GOLD)
