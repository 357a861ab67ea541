"""Pins for the CPU oracle (oracle/phobic_oracle.c) -- CPU only.

Each check is independent of the oracle's own code path:
  * XXH3-128: digests from python xxhash (tests/golden/xxh3_vectors.json)
  * fulcrums / B: tables printed by the reference's own mphf_config.h
    (tests/golden/fulcrums.npz, and oracle/_ref/ref_config when built)
  * hi32 multiply: exact Python integers
  * mix32: inverted with the inverse of the reference's odd multiplier
  * bucket mapping: the 64-bit host formulation of mphf.hpp:42-49
  * bucket order: an independent (-size, -index) sort
  * search: brute-force minimality over every smaller pilot value
  * every build: bijection onto [0, n) (example.cpp:43-59)
"""
import json
import os
import random
import subprocess

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _vector_bytes(length, data_seed):
    words = po.splitmix64(data_seed, (length + 7) // 8)
    return words.astype("<u8").tobytes()[:length]


def test_xxh3_golden_bytes():
    g = json.load(open(os.path.join(GOLD, "xxh3_vectors.json")))
    assert len(g["bytes"]) > 1000
    for v in g["bytes"]:
        data = _vector_bytes(v["len"], v["data_seed"])
        assert po.xxh3_128(data, int(v["seed"])) == int(v["digest"], 16), (v["len"], v["seed"])


def test_xxh3_golden_u64_key_split():
    g = json.load(open(os.path.join(GOLD, "xxh3_vectors.json")))
    for v in g["u64"]:
        k = np.array([int(v["key"])], dtype=np.uint64)
        key = po.hash_u64(k, v["seed"])[0]
        d = int(v["digest"], 16)
        lo, hi = d & (2**64 - 1), d >> 64
        # hasher.hpp:36-37: {lo32(low64), hi32(low64), lo32(high64), hi32(high64)}
        assert [int(x) for x in key] == [lo & 0xFFFFFFFF, lo >> 32, hi & 0xFFFFFFFF, hi >> 32]


def test_xxh3_known_answer_survey():
    # XXH3_128(le64(1), seed 0) = 0xbdc94bce2eda264d_bc08dc21994df8a2
    key = po.hash_u64(np.array([1], dtype=np.uint64), 0)[0]
    assert [int(x) for x in key] == [0x994DF8A2, 0xBC08DC21, 0x2EDA264D, 0xBDC94BCE]


def test_xxh3_live_xxhash():
    xxhash = pytest.importorskip("xxhash")
    rng = random.Random(11)
    for _ in range(300):
        ln = rng.choice([rng.randint(0, 300), rng.randint(0, 3000)])
        seed = rng.getrandbits(64)
        data = bytes(rng.getrandbits(8) for _ in range(ln))
        assert po.xxh3_128(data, seed) == xxhash.xxh3_128_intdigest(data, seed)


def test_hash_bytes_matches_single():
    rng = np.random.default_rng(2)
    lens = rng.integers(0, 100, 500)
    offs = np.zeros(501, np.uint64)
    offs[1:] = np.cumsum(lens)
    data = rng.integers(0, 256, int(offs[-1]), dtype=np.uint8)
    keys = po.hash_bytes(data, offs, 7)
    for i in range(0, 500, 37):
        d = po.xxh3_128(data[int(offs[i]):int(offs[i + 1])].tobytes(), 7)
        lo, hi = d & (2**64 - 1), d >> 64
        assert [int(x) for x in keys[i]] == [lo & 0xFFFFFFFF, lo >> 32, hi & 0xFFFFFFFF, hi >> 32]


def test_fulcrums_match_reference_tables():
    g = np.load(os.path.join(GOLD, "fulcrums.npz"))
    names = [k for k in g.files if k.startswith("f_")]
    assert len(names) >= 6
    for k in names:
        lam, P = k[2:].split("_")
        lam, P = float(lam), int(P)
        assert np.array_equal(po.fulcrums(lam, P), g[k]), k
        B, maxsize = g["b_" + k[2:]]
        assert po.sizes(1, lam, P)[1] == B
        assert maxsize == P + P // 2


def test_fulcrums_match_reference_binary():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_config")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/ref_config not built (needs the reference tree)")
    rng = random.Random(5)
    for _ in range(25):
        lam = round(rng.uniform(1.5, 12.0), 3)
        P = rng.randint(64, 8192)
        out = subprocess.run([exe, repr(lam), str(P)], capture_output=True, text=True, check=True).stdout.split("\n")
        B = int(out[0].split()[0])
        ref = np.array(list(map(int, out[1].split())), dtype=np.uint32)
        assert np.array_equal(po.fulcrums(lam, P), ref), (lam, P)
        assert po.sizes(1, lam, P)[1] == B


def test_sizes_of_configs():
    # B = ceil(P / lambda) for the benchmark configurations
    assert po.sizes(1_000_000, 7.0, 2048) == (489, 293)
    assert po.sizes(100_000_000, 8.0, 2500) == (40000, 313)
    assert po.sizes(1_000_000_000, 7.5, 2048) == (488282, 274)
    assert po.sizes(200_000_000, 6.0, 2048) == (97657, 342)


def test_hi32_multiply_exact():
    lib = po.load()
    rng = random.Random(3)
    cases = [(0, 0), (0xFFFFFFFF, 0xFFFFFFFF), (1, 0xFFFFFFFF), (0x80000000, 2)]
    cases += [(rng.getrandbits(32), rng.getrandbits(32)) for _ in range(20000)]
    for a, b in cases:
        assert lib.orc_hi32_mul(a, b) == (a * b) >> 32


def _mix_inverse(h, pilot):
    inv = pow(0x45D9F3B, -1, 2**32)

    def unshift(x):  # x ^ (x >> 16) is its own inverse on 32 bits
        return x ^ (x >> 16)

    x = unshift(h)
    x = unshift((x * inv) & 0xFFFFFFFF)
    x = unshift((x * inv) & 0xFFFFFFFF)
    return x ^ pilot


def test_mix32_inverts_with_reference_multiplier():
    lib = po.load()
    rng = random.Random(4)
    assert lib.orc_mix32(0, 0) == 0
    for _ in range(20000):
        k, p = rng.getrandbits(32), rng.getrandbits(32)
        assert _mix_inverse(lib.orc_mix32(k, p), p) == k


def _get_bucket_host(x, f):
    # mphf.hpp:42-49, 64-bit formulation
    z = x * 2047
    idx, part = z >> 32, z & 0xFFFFFFFF
    v1 = (int(f[idx]) * part) >> 32
    v2 = (int(f[idx + 1]) * (0xFFFFFFFF - part)) >> 32
    return (v1 + v2) >> 16


def test_bucket_of_matches_host_formulation():
    lib = po.load()
    rng = random.Random(6)
    for lam, P in [(7.0, 2048), (8.0, 2500), (3.0, 64)]:
        f = po.fulcrums(lam, P)
        B = po.sizes(1, lam, P)[1]
        xs = [0, 1, 0xFFFFFFFF, 0x80000000] + [rng.getrandbits(32) for _ in range(20000)]
        for x in xs:
            b = lib.orc_bucket_of(x, f.ctypes.data)
            assert b == _get_bucket_host(x, f)
            assert b < B


def _keys(n, seed=1, hseed=0):
    return po.hash_u64(po.splitmix64(seed, n), hseed)


def _check_bijection(hk, pil, off, lam, P):
    q = po.query_keys(hk, pil, off, lam, P)
    n = hk.shape[0]
    assert q.size == n
    if n:
        assert int(q.max()) < n
        assert np.unique(q).size == n


def test_bucket_order_is_size_desc_index_desc():
    lam, P = 7.0, 2048
    hk = _keys(30000, seed=8)
    st, pil, off, bs, order = po.build_keys(hk, lam, P, intermediates=True)
    assert st == 0
    np_, B = po.sizes(hk.shape[0], lam, P)
    for p in range(np_):
        sizes = bs[p * B:(p + 1) * B]
        nonempty = [b for b in range(B) if sizes[b] > 0]
        expect = sorted(nonempty, key=lambda b: (-int(sizes[b]), -b))
        got = list(order[p * B:p * B + len(expect)])
        assert got == expect
        assert all(order[p * B + len(expect):(p + 1) * B] == 0xFFFFFFFF)
        assert sizes.sum() == off[p + 1] - off[p]


@pytest.mark.parametrize("n,lam,P,alpha", [
    (60000, 7.5, 2048, 0.99), (60000, 7.0, 2048, 1.0), (40000, 3.0, 500, 1.0), (50000, 12.0, 4000, 0.9),
    (30000, 1.0, 256, 1.0),
])
def test_window0_buckets_lead_the_processing_order(n, lam, P, alpha):
    """The search kernel gives the buckets that try offsets 0..63 first (at
    least 6 keys, started while the table is at most half full) a loop of
    their own over ranks [0, r_w0), r_w0 counted from the starts.  That is
    exact only if those buckets are a leading run of the processing order:
    checked here on the oracle's order for every partition."""
    hk = _keys(n, seed=31)
    _, _, off, bs, order = po.build_keys(hk, lam, P, intermediates=True)
    np_, B = po.sizes(n, lam, P)
    lib = po.load()
    a = po.alpha_ppm(alpha)
    leading_runs = 0
    for p in range(np_):
        m = int(lib.orc_table_size(int(off[p + 1] - off[p]), a))
        nb = int((bs[p * B:(p + 1) * B] > 0).sum())
        sizes = bs[p * B + order[p * B:p * B + nb].astype(np.int64)].astype(np.int64)
        starts = np.concatenate(([0], np.cumsum(sizes)[:-1]))
        w0 = (sizes >= 6) & (2 * starts < m)
        k = int(w0.sum())
        assert w0[:k].all() and not w0[k:].any(), p
        leading_runs += k > 0
    if lam >= 7.0:
        assert leading_runs > 0  # the configuration exercises the window loop


def _mix_np(key, pilot):
    key = (key ^ np.uint32(pilot)).astype(np.uint32)
    with np.errstate(over="ignore"):
        key = (((key >> np.uint32(16)) ^ key) * np.uint32(0x45D9F3B)).astype(np.uint32)
        key = (((key >> np.uint32(16)) ^ key) * np.uint32(0x45D9F3B)).astype(np.uint32)
    return (key >> np.uint32(16)) ^ key


@pytest.mark.parametrize("n,lam,P,seed,alpha", [
    (3000, 4.0, 100, 21, 1.0), (2500, 3.0, 64, 22, 1.0), (6000, 7.0, 256, 23, 1.0),
    (6000, 7.0, 600, 24, 0.9),     # tables of m = ceil(ps / 0.9) slots, above 512
    (4096, 7.5, 2048, 25, 0.99),   # the 1e9-key configuration: m crosses 2048
    (5000, 6.0, 700, 26, 0.7),
])
def test_search_minimal_pilots_bruteforce(n, lam, P, seed, alpha):
    """Replays every partition: each stored V is valid and no smaller value
    is, over a table of m slots (m = ps at alpha 1, ceil(ps / alpha) below);
    for alpha < 1 the remap table sends the k-th used slot of [ps, m) to the
    k-th free slot of [0, ps)."""
    hk = _keys(n, seed=seed)
    _, _, _, bs, order = po.build_keys(hk, lam, P, intermediates=True)  # order does not depend on alpha
    st, pil, off, rm, roff = po.build_keys_alpha(hk, lam, P, alpha)
    assert st == 0
    np_, B = po.sizes(n, lam, P)
    f = po.fulcrums(lam, P)
    lib = po.load()
    a = po.alpha_ppm(alpha)
    part = np.array([lib.orc_partition_of(int(k[0]), np_) for k in hk])
    bkt = np.array([lib.orc_bucket_of(int(k[1]), f.ctypes.data) for k in hk])
    for p in range(np_):
        ps = int(off[p + 1] - off[p])
        m = int(lib.orc_table_size(ps, a))
        taken = np.zeros(m, dtype=bool)
        nb = int((bs[p * B:(p + 1) * B] > 0).sum())
        for r in range(nb):
            b = int(order[p * B + r])
            sel = (part == p) & (bkt == b)
            l1, l2 = hk[sel, 2], hk[sel, 3]
            V = int(pil[p + b * np_])
            found = None
            for pilot in range(V // m + 1):
                pos = (_mix_np(l1, _mix_np(l2, pilot)) >> np.uint32(1)) % np.uint32(m)
                if np.unique(pos).size != pos.size:
                    continue
                d = np.arange(m)
                slots = (pos[None, :].astype(np.int64) + d[:, None]) % m
                ok = ~taken[slots].any(axis=1)
                if ok.any():
                    found = pilot * m + int(np.argmax(ok))
                    break
            assert found == V, (p, b, found, V)
            pos = (_mix_np(l1, _mix_np(l2, V // m)) >> np.uint32(1)) % np.uint32(m)
            taken[(pos.astype(np.int64) + V % m) % m] = True
        assert taken.sum() == ps
        if alpha < 1.0:
            used_high = np.nonzero(taken[ps:])[0]
            holes = np.nonzero(~taken[:ps])[0]
            assert used_high.size == holes.size
            want = np.zeros(m - ps, np.uint32)
            want[used_high] = holes
            assert np.array_equal(rm[int(roff[p]):int(roff[p + 1])], want)
    if alpha >= 1.0:
        _check_bijection(hk, pil, off, lam, P)
    else:
        q = po.query_keys_alpha(hk, pil, off, rm, roff, lam, P, alpha)
        assert np.unique(q).size == n and int(q.max()) < n


def test_c1_golden_and_bijection():
    g = np.load(os.path.join(GOLD, "c1_pilots.npz"))
    hk = _keys(1_000_000, seed=1)
    st, pil, off = po.build_keys(hk, 7.0, 2048)
    assert st == 0
    assert np.array_equal(pil, g["pilots"])
    assert np.array_equal(off, g["offsets"])
    _check_bijection(hk, pil, off, 7.0, 2048)


@pytest.mark.parametrize("lam,P", [(8.0, 2500), (7.5, 2048), (6.0, 2048), (4.5, 2500), (12.0, 1024)])
def test_bijection_configs(lam, P):
    hk = _keys(300_000, seed=int(lam * 10) + P)
    st, pil, off = po.build_keys(hk, lam, P)
    assert st == 0
    _check_bijection(hk, pil, off, lam, P)


def test_empty_and_tiny_inputs():
    st, pil, off = po.build_keys(np.zeros((0, 4), np.uint32), 7.0, 2048)
    assert st == 0 and pil.size == 0 and list(off) == [0]
    for n in (1, 2, 7, 100):
        hk = _keys(n, seed=40 + n)
        st, pil, off = po.build_keys(hk, 7.0, 2048)
        assert st == 0 and int(off[-1]) == n
        _check_bijection(hk, pil, off, 7.0, 2048)


def test_duplicate_keys_detected():
    hk = _keys(5000, seed=9)
    hk[4000] = hk[17]
    st, _, _ = po.build_keys(hk, 7.0, 2048)
    assert st == 7


def test_bucket_overflow_detected():
    hk = _keys(4000, seed=10)
    hk[:300, 0] = 5  # partition 0
    hk[:300, 1] = 12345  # one bucket
    st, _, _ = po.build_keys(hk, 7.0, 2048)
    assert st == 6


def test_partition_overflow_detected():
    P = 256
    hk = _keys(4 * P, seed=12)
    hk[: P + P // 2 + 1, 0] = 0  # > 1.5 P keys into partition 0
    st, _, _ = po.build_keys(hk, 7.0, P)
    assert st == 5


def test_compact_vector_known_answer():
    words, w = po.compact_encode(np.array([1, 2, 3], np.uint64))
    assert w == 2 and list(words) == [1 | (2 << 2) | (3 << 4), 0]
    words, w = po.compact_encode(np.array([0, 0, 0], np.uint64))
    assert w == 1 and words.size == 2 and int(words[0]) == 0


@pytest.mark.parametrize("n,maxv", [(1, 1), (64, 2**20), (1000, 2**31 - 1), (333, 5), (77, 2**63)])
def test_compact_vector_roundtrip(n, maxv):
    rng = np.random.default_rng(n)
    vals = rng.integers(0, maxv, n, dtype=np.uint64, endpoint=True)
    words, w = po.compact_encode(vals)
    assert w == max(1, int(vals.max()).bit_length())
    assert words.size == (n * w + 63) // 64 + 1
    for i in range(n):
        assert po.compact_access(words, w, i) == int(vals[i])


def _parse_mphf(buf):
    """Decodes the serialised MPHF ("PHOBMPH2") without the oracle: header,
    fulcrums, pilot columns (compact or Rice; Rice values are read by
    scanning the unary bits, and the select index is rebuilt from the one
    positions and compared), partition offsets."""
    v = np.frombuffer(buf, dtype="<u8")
    magic, n, P, np_, B, seed, enc, a_ppm = (int(x) for x in v[:8])
    at = 8
    fl = v[at:at + 1024].view("<u4")
    at += 1024

    def compact(at):
        size, w, mask, nw = (int(x) for x in v[at:at + 4])
        words = np.ascontiguousarray(v[at + 4:at + 4 + nw])
        vals = [po.compact_access(words, w, i) for i in range(size)] if w else [0] * size
        return vals, w, at + 4 + nw

    def rice(at):
        bits, hw = int(v[at]), int(v[at + 1])
        words = [int(x) for x in v[at + 2:at + 2 + hw]]
        at += 2 + hw
        ones = [i for i in range(bits) if (words[i >> 6] >> (i & 63)) & 1]
        count, nb = int(v[at]), int(v[at + 1])
        blocks = [int(x) for x in v[at + 2:at + 2 + nb]]
        at += 2 + nb
        ns = int(v[at])
        packed = v[at + 1:at + 1 + (ns + 3) // 4]
        subs = [(int(packed[i // 4]) >> (16 * (i % 4))) & 0xFFFF for i in range(ns)]
        at += 1 + (ns + 3) // 4
        nsp = int(v[at])
        sparse = [int(x) for x in v[at + 1:at + 1 + nsp]]
        at += 1 + nsp
        low, l, at = compact(at)
        assert count == len(ones) == len(low)
        eb, es, esp = [], [], []
        for b0 in range(0, count, 4096):
            blk = ones[b0:b0 + 4096]
            if blk[-1] - blk[0] < 65536:
                eb.append(blk[0])
                es += [blk[i] - blk[0] for i in range(0, len(blk), 64)]
            else:
                eb.append((-len(esp) - 1) & (2**64 - 1))
                esp += blk
                es += [0xFFFF] * ((len(blk) + 63) // 64)
        assert blocks == eb and subs == es and sparse == esp
        highs = np.diff(np.array([-1] + ones)) - 1
        sparse_entries.append(len(sparse))
        return [(int(h) << l) | x for h, x in zip(highs, low)], l, at

    cols, params, sparse_entries = [], [], []
    for _ in range(B if np_ else 0):
        c, w, at = compact(at) if enc == 0 else rice(at)
        cols.append(c)
        params.append(w)
    inc = int(v[at])
    diffs, _, at = compact(at + 1)
    assert at == v.size
    return dict(magic=magic, n=n, P=P, np=np_, B=B, seed=seed, enc=enc, a_ppm=a_ppm, fl=fl, cols=cols,
                params=params, sparse=sparse_entries, inc=inc, diffs=diffs)


def test_encoded_mphf_layout_decodes():
    lam, P = 7.0, 512
    hk = _keys(20000, seed=30)
    st, pil, off = po.build_keys(hk, lam, P)
    buf = po.encode_mphf(pil, off, lam, P, hk.shape[0], seed=7)
    m = _parse_mphf(buf)
    np_, B, cols, inc, diffs = m["np"], m["B"], m["cols"], m["inc"], m["diffs"]
    assert m["magic"] == 0x3248504D424F4850 and m["n"] == 20000 and m["P"] == P and inc == P
    assert m["seed"] == 7 and m["enc"] == 0 and m["a_ppm"] == 1000000
    assert np.array_equal(m["fl"], po.fulcrums(lam, P))
    for j in range(B):
        assert cols[j] == [int(x) for x in pil[j * np_:(j + 1) * np_]]
    # linear_diff_encoder: value = (|v| << 1) | (v > 0), v = off[i] - i*P
    for i, e in enumerate(diffs):
        s = (2 * (e & 1) - 1) * (e >> 1)
        assert s + i * P == int(off[i])


def test_rice_columns_decode_to_the_pilots():
    lam, P = 7.0, 512
    hk = _keys(20000, seed=31)
    st, pil, off = po.build_keys(hk, lam, P)
    buf, bpk = po.encode_mphf(pil, off, lam, P, hk.shape[0], encoding="rice", with_bits_per_key=True)
    m = _parse_mphf(buf)
    np_ = m["np"]
    assert m["enc"] == 1
    for j in range(m["B"]):
        col = pil[j * np_:(j + 1) * np_].astype(np.uint64)
        assert m["cols"][j] == [int(x) for x in col]
        assert m["params"][j] == po.rice_param(col)
    assert bpk > 0


def test_rice_is_smaller_for_long_columns():
    # ~1000 partitions per column: Rice codes beat fixed-width columns once
    # the per-column index overhead is amortised
    lam, P = 7.0, 512
    hk = _keys(500_000, seed=32)
    st, pil, off = po.build_keys(hk, lam, P)
    _, rice_bpk = po.encode_mphf(pil, off, lam, P, hk.shape[0], encoding="rice", with_bits_per_key=True)
    _, fast_bpk = po.encode_mphf(pil, off, lam, P, hk.shape[0], with_bits_per_key=True)
    assert 0 < rice_bpk < fast_bpk


def test_rice_sparse_select_blocks():
    # two maximal values among zeros: their unary parts (2^32 >> l with
    # l about 14) make their blocks of 4096 ones span more than 2^16 bits
    n = 200_000
    vals = np.zeros(n, np.uint64)
    vals[[5000, 150000]] = 2**32 - 1
    pil = vals.astype(np.uint32)  # one bucket column over n partitions
    off = np.arange(n + 1, dtype=np.uint32)
    buf = po.encode_mphf(pil, off, 1.0, 1, n, encoding="rice")
    m = _parse_mphf(buf)
    assert m["B"] == 1 and m["cols"][0] == [int(x) for x in vals]
    assert m["sparse"][0] == 2 * 4096


@pytest.mark.parametrize("mean", [0.3, 1.0, 4.0, 37.0, 500.0, 70000.0])
def test_rice_parameter_near_optimal_for_geometric_values(mean):
    """On geometric samples the chosen l gives a code length within 1% of
    the best l found by trying every l (bits = n*(l+1) + sum(v >> l))."""
    rng = np.random.default_rng(int(mean * 10))
    vals = rng.geometric(1.0 / (mean + 1.0), 200_000).astype(np.uint64) - np.uint64(1)
    cost = [len(vals) * (l + 1) + int((vals >> np.uint64(l)).sum()) for l in range(40)]
    l = po.rice_param(vals)
    assert cost[l] <= 1.01 * min(cost)
    assert po.rice_param(np.zeros(100, np.uint64)) == 0


@pytest.mark.parametrize("n,lam,P,alpha,kseed", [
    (300_000, 7.5, 2048, 0.99, 3),
    (200_000, 4.0, 500, 1.0, 4),
    (150_000, 7.0, 1024, 0.7, 5),
])
def test_partition_subsets_equal_full_build(n, lam, P, alpha, kseed):
    """The streaming partition builder (used at sizes the full oracle build
    cannot hold) equals orc_build_keys_ex row by row, for partitions listed
    in any order, and its scan gives the full build's counts and largest
    bucket per partition."""
    hk = po.hash_u64(po.splitmix64(kseed, n), 0)
    st, pil, off, rm, roff = po.build_keys_alpha(hk, lam, P, alpha)
    assert st == 0
    np_, B = po.sizes(n, lam, P)
    parts = np.random.default_rng(kseed).permutation(np_)[: max(1, np_ // 3)].astype(np.uint32)
    st2, pil2, cnt2, rm2 = po.build_splitmix_parts(kseed, n, 0, lam, P, alpha, parts)
    assert st2 == 0
    assert np.array_equal(pil2, pil.reshape(B, np_).T[parts])
    assert np.array_equal(cnt2, np.diff(off.astype(np.int64))[parts])
    if alpha < 1.0:
        assert np.array_equal(rm2, np.concatenate([rm[roff[p]:roff[p + 1]] for p in parts]))
    counts, maxb = po.scan_splitmix(kseed, n, 0, lam, P)
    assert np.array_equal(counts, np.diff(off.astype(np.int64)))
    _, _, _, bs, _ = po.build_keys(hk, lam, P, intermediates=True)
    assert np.array_equal(maxb, bs.reshape(np_, B).max(axis=1))


def test_partition_subset_errors():
    # a partition listed twice or out of range is refused
    assert po.build_splitmix_parts(1, 10_000, 0, 7.0, 2048, 1.0, [0, 0])[0] == 1
    assert po.build_splitmix_parts(1, 10_000, 0, 7.0, 2048, 1.0, [5])[0] == 1


def test_custom_bucketer_tables_match_reference():
    """MPHFconfig(Bucketer*, lambda, P) with the reference's UniformBucketer and
    SkewBucketer (tables printed by the reference-built ref_config, committed
    in fulcrums.npz): B = round(P / lambda) and the same 2048 entries."""
    g = np.load(os.path.join(GOLD, "fulcrums.npz"))
    for kind in ("uniform", "skew"):
        for lam, P in ((7.0, 2048), (4.5, 1000)):
            B, f = po.bucketer_fulcrums(kind, lam, P)
            assert B == int(g[f"{kind}b_{lam}_{P}"][0])
            assert np.array_equal(f, g[f"{kind}_{lam}_{P}"]), (kind, lam, P)


def test_custom_table_build_is_a_bijection():
    """A build with a skew table differs from the OptBucketer build and is a
    bijection (the oracle's query uses the same table)."""
    n, P = 200_000, 2048
    B, f = po.bucketer_fulcrums("skew", 7.0, P)
    hk = po.hash_u64(po.splitmix64(13, n), 0)
    st, pil, off, _, _ = po.build_keys_table(hk, P, B, f)
    assert st == 0
    lib = po.load()
    q = np.zeros(n, np.uint32)
    lib.orc_query_keys(hk.ctypes.data, n, pil.ctypes.data, off.ctypes.data, off.size - 1, B, f.ctypes.data,
                       q.ctypes.data)
    assert np.unique(q).size == n and int(q.max()) < n
    # an entry above B << 16 would map keys past the last bucket: refused
    bad = f.copy()
    bad[1000] = (B << 16) + 1
    assert po.build_keys_table(hk, P, B, bad)[0] == 1


def test_device_outputs_keeps_its_buffers_alive():
    """phobic_hip.device_outputs holds the buffer objects it was given, so a
    buffer created inside the call cannot be freed while a build still writes
    to it (a temporary freed mid-build once faulted a concurrent build)."""
    import weakref

    import phobic_hip as ph

    class Buf:
        def __init__(self, ptr):
            self.ptr = ptr

    bufs = [Buf(0x1000), Buf(0x2000), Buf(0x3000), Buf(0x4000)]
    refs = [weakref.ref(b) for b in bufs]
    out = ph.device_outputs(*bufs)
    del bufs
    assert all(r() is not None for r in refs)
    assert (out.pilots, out.offsets, out.remap, out.remap_offsets) == (0x1000, 0x2000, 0x3000, 0x4000)
    del out
    assert all(r() is None for r in refs)
