"""Parity of the HIP construction path with the oracle (bit-exact pilots and
partition offsets), through the C ABI, on an MI355X."""
import os

import numpy as np
import pytest

import hip_rt
import phobic_hip as ph
import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_u64(keys, lam, P, seed):
    hk = po.hash_u64(keys, seed)
    st, pil, off = po.build_keys(hk, lam, P)
    return st, pil, off, hk


def assert_bijection(hk, pil, off, lam, P):
    q = po.query_keys(hk, pil, off, lam, P)
    n = hk.shape[0]
    if n:
        assert int(q.max()) < n and np.unique(q).size == n


def test_c1_bit_exact_vs_golden(ctx):
    g = np.load(os.path.join(GOLD, "c1_pilots.npz"))
    keys = po.splitmix64(1, 1_000_000)
    r = ctx.build_u64(keys, ph.make_config(7.0, 2048, 0))
    assert np.array_equal(r.offsets, g["offsets"])
    assert np.array_equal(r.pilots, g["pilots"])


@pytest.mark.parametrize("n,lam,P,seed,kseed", [
    (1_000_000, 7.0, 2048, 42, 1),   # 1M-key configuration with hash seed 42
    (2_000_000, 8.0, 2500, 0, 2),    # 100M-key configuration's lambda / P
    (1_000_000, 7.5, 2048, 0, 3),    # 1B-key configuration's lambda / P
    (1_000_000, 6.0, 2048, 0, 4),    # string configuration's lambda / P
    (500_000, 4.5, 2500, 0, 5),      # reference example.cpp configuration
    (200_000, 12.0, 1024, 7, 6),
    (5_000, 3.0, 64, 0, 7),
    (300_000, 8.0, 5000, 0, 8),      # tables above 4096 slots: four sub-windows per lane
    (200_000, 4.0, 9000, 0, 9),      # tables above 8192 slots: several scan passes
    (300_000, 1.0, 2048, 0, 10),     # lambda 1: B = P buckets per partition, mostly one key
    (300_000, 10.0, 2048, 0, 11),    # lambda 10: 205 buckets, the first ones well above 64 keys
])
def test_u64_bit_exact_vs_oracle(ctx, n, lam, P, seed, kseed):
    keys = po.splitmix64(kseed, n)
    r = ctx.build_u64(keys, ph.make_config(lam, P, seed))
    st, pil, off, hk = oracle_u64(keys, lam, P, seed)
    assert st == 0
    assert np.array_equal(r.offsets, off)
    assert np.array_equal(r.pilots, pil)
    assert_bijection(hk, r.pilots, r.offsets, lam, P)


def test_split_with_more_parents_than_the_pipeline_takes(ctx):
    """Fan-out 16 at 19,532 partitions (5M keys, P 256): four routing levels,
    the last one splitting 1,221 parent bins, more than the pipelined split
    kernel takes (1024), so it runs the per-tile split; the levels before it
    run the pipelined kernels (raw keys to raw keys).  Bit-exact against the
    oracle."""
    keys = po.splitmix64(31, 5_000_000)
    lam, P = 4.0, 256
    saved = ctx.tuning()
    try:
        ctx.set_tuning(**{**saved, "max_fanout": 16})
        r = ctx.build_u64(keys, ph.make_config(lam, P))
    finally:
        ctx.set_tuning(**saved)
    st, pil, off, hk = oracle_u64(keys, lam, P, 0)
    assert st == 0 and r.np_ == 19532
    assert np.array_equal(r.offsets, off)
    assert np.array_equal(r.pilots, pil)


def random_strings(n, lo, hi, seed, printable=True):
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi + 1, n)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    if printable:
        data = rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8)
    else:
        data = rng.integers(0, 256, int(offs[-1]), dtype=np.uint8)
    return data, offs


def test_strings_bit_exact(ctx):
    data, offs = random_strings(300_000, 8, 64, 3)
    r = ctx.build_bytes(data, offs, ph.make_config(6.0, 2048, 0))
    hk = po.hash_bytes(data, offs, 0)
    st, pil, off = po.build_keys(hk, 6.0, 2048)
    assert st == 0
    assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)


@pytest.mark.parametrize("seed", [0, 42, 2**64 - 1])
def test_strings_every_xxh3_length_class(ctx, seed):
    # lengths 0..300 cover the 0, 1-3, 4-8, 9-16, 17-128, 129-240 and long paths
    rng = np.random.default_rng(seed % 1000)
    lens = np.concatenate([np.arange(0, 301), [511, 1024, 1025, 4096, 5000]])
    lens = np.repeat(lens, 20)
    offs = np.zeros(lens.size + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    data = rng.integers(0, 256, int(offs[-1]), dtype=np.uint8)
    # distinct strings: the first 8 bytes of each carry its index (len >= 8);
    # short strings are deduplicated below
    for i in range(lens.size):
        if lens[i] >= 8:
            data[int(offs[i]):int(offs[i]) + 8] = np.frombuffer(np.uint64(i).tobytes(), np.uint8)
    strs = {}
    for i in range(lens.size):
        strs.setdefault(data[int(offs[i]):int(offs[i + 1])].tobytes(), i)
    keep = sorted(strs.values())
    lens2 = lens[keep]
    offs2 = np.zeros(len(keep) + 1, np.uint64)
    offs2[1:] = np.cumsum(lens2)
    data2 = np.concatenate([data[int(offs[i]):int(offs[i + 1])] for i in keep])
    cfg = ph.make_config(5.0, 256, seed)
    r = ctx.build_bytes(data2, offs2, cfg)
    hk = po.hash_bytes(data2, offs2, seed)
    st, pil, off = po.build_keys(hk, 5.0, 256)
    assert st == 0
    assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)


def test_prehashed_keys_bit_exact(ctx):
    hk = po.hash_u64(po.splitmix64(8, 700_000), 0)
    r = ctx.build_keys(hk, ph.make_config(7.0, 2048))
    st, pil, off = po.build_keys(hk, 7.0, 2048)
    assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)


def test_empty_and_tiny(ctx):
    r = ctx.build_u64(np.zeros(0, np.uint64), ph.make_config(7.0, 2048))
    assert r.pilots.size == 0 and list(r.offsets) == [0]
    for n in (1, 2, 3, 64, 2049):
        keys = po.splitmix64(100 + n, n)
        r = ctx.build_u64(keys, ph.make_config(7.0, 2048))
        st, pil, off, hk = oracle_u64(keys, 7.0, 2048, 0)
        assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)


def test_duplicate_keys_error(ctx):
    keys = po.splitmix64(9, 10_000)
    keys[5000] = keys[17]
    with pytest.raises(ph.PhobicError) as e:
        ctx.build_u64(keys, ph.make_config(7.0, 2048))
    assert e.value.status == 7


@pytest.mark.parametrize("size", [65, 150, 200])
def test_large_bucket_matches_oracle(ctx, size):
    """One bucket far above 64 keys (the chunked collision path): pilots and
    offsets equal the oracle's."""
    hk = po.hash_u64(po.splitmix64(11 + size, 6000), 0)
    hk[:size, 0] = 5
    hk[:size, 1] = 0xF0000000  # a late bucket: few keys of its own
    r = ctx.build_keys(hk, ph.make_config(7.0, 2048))
    st, pil, off = po.build_keys(hk, 7.0, 2048)
    assert st == 0
    assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)


def test_pilot_overflow_error(ctx):
    """250 keys in one bucket of a ~2300-slot table need about 10^6 pilots:
    V = d + tm * pilot passes 2^31 first, which both report."""
    hk = po.hash_u64(po.splitmix64(261, 6000), 0)
    hk[:250, 0] = 5
    hk[:250, 1] = 0xF0000000
    with pytest.raises(ph.PhobicError) as e:
        ctx.build_keys(hk, ph.make_config(7.0, 2048))
    assert e.value.status == 8
    assert po.build_keys(hk, 7.0, 2048)[0] == 8


def test_duplicate_keys_in_large_bucket_error(ctx):
    hk = po.hash_u64(po.splitmix64(12, 6000), 0)
    hk[:100, 0] = 5
    hk[:100, 1] = 12345
    hk[40, 2:] = hk[7, 2:]
    with pytest.raises(ph.PhobicError) as e:
        ctx.build_keys(hk, ph.make_config(7.0, 2048))
    assert e.value.status == 7
    assert po.build_keys(hk, 7.0, 2048)[0] == 7


def test_bucket_overflow_error(ctx):
    hk = po.hash_u64(po.splitmix64(10, 4000), 0)
    hk[:300, 0] = 5
    hk[:300, 1] = 12345
    with pytest.raises(ph.PhobicError) as e:
        ctx.build_keys(hk, ph.make_config(7.0, 2048))
    assert e.value.status == 6
    assert po.build_keys(hk, 7.0, 2048)[0] == 6


def test_partition_overflow_error(ctx):
    P = 256
    hk = po.hash_u64(po.splitmix64(12, 4 * P), 0)
    hk[: P + P // 2 + 1, 0] = 0
    with pytest.raises(ph.PhobicError) as e:
        ctx.build_keys(hk, ph.make_config(7.0, P))
    assert e.value.status == 5


def test_context_reuse_after_error(ctx):
    keys = po.splitmix64(13, 50_000)
    r = ctx.build_u64(keys, ph.make_config(7.0, 2048))
    st, pil, off, hk = oracle_u64(keys, 7.0, 2048, 0)
    assert np.array_equal(r.pilots, pil)


def test_deterministic_repeat(ctx):
    keys = po.splitmix64(14, 2_000_000)
    cfg = ph.make_config(7.0, 2048)
    a = ctx.build_u64(keys, cfg)
    b = ctx.build_u64(keys, cfg)
    assert np.array_equal(a.pilots, b.pilots) and np.array_equal(a.offsets, b.offsets)


def test_device_generator_matches_numpy(ctx):
    n = 1_000_003
    d = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(77, n, d)
    hip_rt.synchronize()
    assert np.array_equal(d.to_numpy(np.uint64), po.splitmix64(77, n))


def test_device_build_matches_host_build(ctx):
    n, lam, P = 3_000_000, 8.0, 2500
    cfg = ph.make_config(lam, P)
    np_, B = ph.sizes(cfg, n)
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(2, n, dk)
    dp, do, ds = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    ctx.build_u64_device(dk, n, cfg, ph.device_outputs(dp, do), ds)
    hip_rt.synchronize()
    assert int(ds.to_numpy(np.uint32)[0]) == 0
    r = ctx.build_u64(po.splitmix64(2, n), cfg)
    assert np.array_equal(dp.to_numpy(np.uint32), r.pilots)
    assert np.array_equal(do.to_numpy(np.uint32), r.offsets)


def test_large_host_build_matches_device_build(ctx):
    """Host-buffer builds large enough that keys and pilots move through the
    pinned staging chunks (16 chunks of 64 MB in, 2 out) equal the device
    build of the same keys, remap table included."""
    n, lam, P, alpha = 130_000_000, 7.5, 2048, 0.99
    cfg = ph.make_config(lam, P, 0, alpha)
    np_, B = ph.sizes(cfg, n)
    rcap = ph.remap_capacity(cfg, n)
    assert n * 8 > 15 * (64 << 20) and np_ * B * 4 > 64 << 20  # several chunks both ways
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(9, n, dk)
    dp, do, ds = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    drm, dro = hip_rt.DeviceBuffer(rcap * 4), hip_rt.DeviceBuffer((np_ + 1) * 4)
    ctx.build_u64_device(dk, n, cfg, ph.device_outputs(dp, do, drm, dro), ds)
    hip_rt.synchronize()
    assert int(ds.to_numpy(np.uint32)[0]) == 0
    keys = dk.to_numpy(np.uint64)
    for _ in range(2):  # the second call reuses the pinned chunks
        r = ctx.build_u64(keys, cfg)
        assert np.array_equal(dp.to_numpy(np.uint32), r.pilots)
        assert np.array_equal(do.to_numpy(np.uint32), r.offsets)
        roff = dro.to_numpy(np.uint32)
        assert np.array_equal(roff, r.remap_offsets)
        assert np.array_equal(drm.to_numpy(np.uint32, int(roff[-1])), r.remap)


def test_large_host_string_build_matches_device_build(ctx):
    """Strings whose bytes and offsets both move through the staging chunks."""
    rng = np.random.default_rng(21)
    n = 3_000_000
    lens = rng.integers(8, 65, n)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    assert int(offs[-1]) > 64 << 20 and (n + 1) * 8 > 16 << 20
    d_offs = hip_rt.DeviceBuffer.from_numpy(offs)
    d_bytes = hip_rt.DeviceBuffer(int(offs[-1]))
    ph.gen_strings_device(79, n, d_offs, d_bytes)
    hip_rt.synchronize()
    data = d_bytes.to_numpy(np.uint8, int(offs[-1]))
    cfg = ph.make_config(6.0, 2048)
    np_, B = ph.sizes(cfg, n)
    dp, do, ds = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    ctx.build_bytes_device(d_bytes, d_offs, n, cfg, ph.device_outputs(dp, do), ds)
    hip_rt.synchronize()
    assert int(ds.to_numpy(np.uint32)[0]) == 0
    r = ctx.build_bytes(data, offs, cfg)
    assert np.array_equal(dp.to_numpy(np.uint32), r.pilots)
    assert np.array_equal(do.to_numpy(np.uint32), r.offsets)


def test_device_bijection_100m(ctx):
    """100M distinct keys (splitmix64 seed 2), lambda 8, P 2500: every key maps
    to a distinct slot of [0, n), checked on the device."""
    n, lam, P = 100_000_000, 8.0, 2500
    cfg = ph.make_config(lam, P)
    np_, B = ph.sizes(cfg, n)
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(2, n, dk)
    dp, do, ds = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    mphf = ph.device_outputs(dp, do)
    ctx.build_u64_device(dk, n, cfg, mphf, ds)
    hip_rt.synchronize()
    assert int(ds.to_numpy(np.uint32)[0]) == 0
    off = do.to_numpy(np.uint32)
    assert int(off[-1]) == n and np.all(np.diff(off.astype(np.int64)) <= P + P // 2)
    out = hip_rt.DeviceBuffer(n * 4)
    ctx.eval_u64_device(dk, n, cfg, np_, mphf, out)
    bm, bad = hip_rt.DeviceBuffer((n + 31) // 32 * 4), hip_rt.DeviceBuffer(8)
    ctx.check_bijection_device(out, n, n, bm, bad)
    hip_rt.synchronize()
    assert int(bad.to_numpy(np.uint64)[0]) == 0


@pytest.mark.parametrize("n,lam,P,alpha,kseed", [
    (1_000_000, 7.5, 2048, 0.99, 3),   # the 1B-key configuration's lambda / alpha
    (400_000, 6.0, 2048, 0.99, 4),
    (300_000, 8.0, 2500, 0.95, 5),
    (100_000, 7.0, 1024, 0.5, 6),
    (200_000, 7.0, 2500, 0.55, 7),     # alpha tables above 4096 slots
])
def test_alpha_remap_bit_exact(ctx, n, lam, P, alpha, kseed):
    keys = po.splitmix64(kseed, n)
    r = ctx.build_u64(keys, ph.make_config(lam, P, 0, alpha))
    hk = po.hash_u64(keys, 0)
    st, pil, off, rm, roff = po.build_keys_alpha(hk, lam, P, alpha)
    assert st == 0
    assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)
    assert np.array_equal(r.remap_offsets, roff) and np.array_equal(r.remap, rm)
    q = po.query_keys_alpha(hk, r.pilots, r.offsets, r.remap, r.remap_offsets, lam, P, alpha)
    assert int(q.max()) < n and np.unique(q).size == n


def test_alpha_device_eval_bijection(ctx):
    n, lam, P, alpha = 5_000_000, 7.5, 2048, 0.99
    cfg = ph.make_config(lam, P, 0, alpha)
    np_, B = ph.sizes(cfg, n)
    rcap = ph.remap_capacity(cfg, n)
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(31, n, dk)
    dp, do = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4)
    drm, dro, ds = hip_rt.DeviceBuffer(rcap * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    mphf = ph.device_outputs(dp, do, drm, dro)
    ctx.build_u64_device(dk, n, cfg, mphf, ds)
    out = hip_rt.DeviceBuffer(n * 4)
    ctx.eval_u64_device(dk, n, cfg, np_, mphf, out)
    bm, bad = hip_rt.DeviceBuffer((n + 31) // 32 * 4), hip_rt.DeviceBuffer(8)
    ctx.check_bijection_device(out, n, n, bm, bad)
    hip_rt.synchronize()
    assert int(ds.to_numpy(np.uint32)[0]) == 0
    assert int(bad.to_numpy(np.uint64)[0]) == 0


@pytest.mark.parametrize("n,lam,P,alpha,kseed", [
    (300_000, 8.0, 2500, 1.0, 11),
    (300_000, 8.0, 5000, 1.0, 12),
    (200_000, 4.0, 9000, 1.0, 13),
    (200_000, 7.0, 2500, 0.55, 14),
])
def test_search_modes_bit_exact(ctx, n, lam, P, alpha, kseed):
    """Every pilot-search variant (one or several pilots per hash round,
    wide or 2048-offset scan passes) gives the oracle's pilots."""
    keys = po.splitmix64(kseed, n)
    cfg = ph.make_config(lam, P, 0, alpha)
    hk = po.hash_u64(keys, 0)
    if alpha < 1.0:
        st, opil, ooff, orm, _ = po.build_keys_alpha(hk, lam, P, alpha)
    else:
        st, opil, ooff = po.build_keys(hk, lam, P)
    assert st == 0
    saved = ctx.tuning()
    try:
        for mode in (0, 1, 2, 3):
            ctx.set_tuning(**{**saved, "search_mode": mode})
            r = ctx.build_u64(keys, cfg)
            assert np.array_equal(r.pilots, opil), mode
            assert np.array_equal(r.offsets, ooff), mode
    finally:
        ctx.set_tuning(**saved)


@pytest.mark.parametrize("alpha", [1.0, 0.97])
def test_offsets_span_several_scan_tiles(ctx, alpha):
    """20M keys at P = 2048 give 9766 partitions, three tiles of the offsets
    scan and three routing levels: offsets are the exclusive scan of the
    partition counts and remap offsets that of table_size(count) - count;
    pilots and remap entries equal the oracle's."""
    n, lam, P = 20_000_000, 7.5, 2048
    keys = po.splitmix64(31, n)
    r = ctx.build_u64(keys, ph.make_config(lam, P, 0, alpha))
    hk = po.hash_u64(keys, 0)
    np_ = (n + P - 1) // P
    part = (hk[:, 0].astype(np.uint64) * np.uint64(np_)) >> np.uint64(32)
    cnt = np.bincount(part.astype(np.int64), minlength=np_)
    off = np.zeros(np_ + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    assert np.array_equal(r.offsets.astype(np.int64), off)
    if alpha < 1.0:
        a = po.alpha_ppm(alpha)
        lib = po.load()
        extra = np.array([lib.orc_table_size(int(c), a) - int(c) for c in cnt], np.int64)
        roff = np.zeros(np_ + 1, np.int64)
        roff[1:] = np.cumsum(extra)
        assert np.array_equal(r.remap_offsets.astype(np.int64), roff)
        st, pil, ooff, rm, oroff = po.build_keys_alpha(hk, lam, P, alpha)
        assert st == 0 and np.array_equal(r.remap, rm)
    else:
        st, pil, ooff = po.build_keys(hk, lam, P)
        assert st == 0
    assert np.array_equal(r.pilots, pil)


def gen_strings_ref(seed, offs):
    """numpy restatement of phobic_gen_strings_device for small inputs."""
    n = offs.size - 1
    out = np.zeros(int(offs[-1]), np.uint8)
    M = (1 << 64) - 1
    for i in range(n):
        a, ln = int(offs[i]), int(offs[i + 1] - offs[i])
        state = seed ^ ((i * 0xD1B54A32D192ED03) & M)
        r = 0
        for j in range(ln):
            if j % 8 == 0:
                state = (state + 0x9E3779B97F4A7C15) & M
                z = state
                z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
                z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
                r = z ^ (z >> 31)
            out[a + j] = 0x20 + (r & 0xFF) % 95
            r >>= 8
    return out


def test_device_string_generator_and_build(ctx):
    """Device-generated strings equal their numpy restatement, and a build on
    them (lengths 8..64, lambda 6) equals the oracle's."""
    rng = np.random.default_rng(12)
    n = 2000
    lens = rng.integers(8, 65, n)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    d_offs = hip_rt.DeviceBuffer.from_numpy(offs)
    d_bytes = hip_rt.DeviceBuffer(int(offs[-1]))
    ph.gen_strings_device(77, n, d_offs, d_bytes)
    hip_rt.synchronize()
    got = d_bytes.to_numpy(np.uint8, int(offs[-1]))
    assert np.array_equal(got, gen_strings_ref(77, offs))
    # a larger set for the build
    n = 300_000
    lens = rng.integers(8, 65, n)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    d_offs = hip_rt.DeviceBuffer.from_numpy(offs)
    d_bytes = hip_rt.DeviceBuffer(int(offs[-1]))
    ph.gen_strings_device(78, n, d_offs, d_bytes)
    hip_rt.synchronize()
    data = d_bytes.to_numpy(np.uint8, int(offs[-1]))
    r = ctx.build_bytes(data, offs, ph.make_config(6.0, 2048))
    st, opil, ooff = po.build_keys(po.hash_bytes(data, offs, 0), 6.0, 2048)
    assert st == 0
    assert np.array_equal(r.offsets, ooff) and np.array_equal(r.pilots, opil)


@pytest.mark.parametrize("kind,lam,P,alpha", [
    ("uniform", 7.0, 2048, 1.0),
    ("skew", 4.5, 1000, 1.0),
    ("skew", 7.0, 2048, 0.99),
])
def test_custom_bucketer_table_bit_exact(ctx, kind, lam, P, alpha):
    """MPHFconfig(Bucketer*, lambda, P) (mphf_config.h:34-42): a custom B and
    fulcrum table give the oracle's pilots, offsets and remap table."""
    n = 400_000
    B, f = po.bucketer_fulcrums(kind, lam, P)
    keys = po.splitmix64(17, n)
    r = ctx.build_u64(keys, ph.make_config_table(B, f, lam, P, 0, alpha))
    assert r.B == B
    st, pil, off, rm, roff = po.build_keys_table(po.hash_u64(keys, 0), P, B, f, alpha)
    assert st == 0
    assert np.array_equal(r.offsets, off) and np.array_equal(r.pilots, pil)
    if alpha < 1.0:
        assert np.array_equal(r.remap_offsets, roff) and np.array_equal(r.remap, rm)
