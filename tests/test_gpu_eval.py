"""Device evaluation (MPHF::operator(), mphf.hpp:77-86) for every key type of
the reference's hasher (hasher.hpp:22-66), through the C ABI: positions equal
the oracle's query of the same keys, bit for bit, and form a bijection."""
import numpy as np
import pytest

import hip_rt
import phobic_hip as ph
import pyoracle as po

pytestmark = pytest.mark.gpu


def device_mphf(r, alpha):
    bufs = [hip_rt.DeviceBuffer.from_numpy(r.pilots), hip_rt.DeviceBuffer.from_numpy(r.offsets)]
    if alpha < 1.0:
        bufs += [hip_rt.DeviceBuffer.from_numpy(r.remap if r.remap.size else np.zeros(1, np.uint32)),
                 hip_rt.DeviceBuffer.from_numpy(r.remap_offsets)]
    return ph.device_outputs(*bufs)


def oracle_positions(hk, r, lam, P, alpha):
    if alpha < 1.0:
        return po.query_keys_alpha(hk, r.pilots, r.offsets, r.remap, r.remap_offsets, lam, P, alpha)
    return po.query_keys(hk, r.pilots, r.offsets, lam, P)


@pytest.mark.parametrize("n,lam,P,alpha", [(2_000_000, 7.5, 2048, 0.99), (1_000_000, 8.0, 2500, 1.0)])
def test_eval_u64_equals_oracle(ctx, n, lam, P, alpha):
    keys = po.splitmix64(41, n)
    cfg = ph.make_config(lam, P, 0, alpha)
    r = ctx.build_u64(keys, cfg)
    d_keys, d_out = hip_rt.DeviceBuffer.from_numpy(keys), hip_rt.DeviceBuffer(n * 4)
    ctx.eval_u64_device(d_keys, n, cfg, r.np_, device_mphf(r, alpha), d_out)
    hip_rt.synchronize()
    got = d_out.to_numpy(np.uint32)
    want = oracle_positions(po.hash_u64(keys, 0), r, lam, P, alpha)
    assert np.array_equal(got, want)
    assert np.unique(got).size == n and int(got.max()) < n


@pytest.mark.parametrize("alpha", [1.0, 0.99])
def test_eval_strings_equals_oracle(ctx, alpha):
    rng = np.random.default_rng(42)
    n, lam, P = 500_000, 6.0, 2048
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(rng.integers(4, 300, n))  # the XXH3 length classes from 4 bytes on
    data = rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8)
    cfg = ph.make_config(lam, P, 0, alpha)
    r = ctx.build_bytes(data, offs, cfg)
    d_b, d_o, d_out = hip_rt.DeviceBuffer.from_numpy(data), hip_rt.DeviceBuffer.from_numpy(offs), hip_rt.DeviceBuffer(n * 4)
    ctx.eval_bytes_device(d_b, d_o, n, cfg, r.np_, device_mphf(r, alpha), d_out)
    hip_rt.synchronize()
    got = d_out.to_numpy(np.uint32)
    want = oracle_positions(po.hash_bytes(data, offs, 0), r, lam, P, alpha)
    assert np.array_equal(got, want)
    assert np.unique(got).size == n


def test_eval_pairs_and_xxhash_keys_equal_oracle(ctx):
    """std::pair<u64, u64> keys (and xxhash-hashed Keys) hash their 16 bytes;
    the MPHF is built from the same 16-byte strings."""
    n, lam, P = 400_000, 7.0, 2048
    pairs = np.stack([po.splitmix64(51, n), po.splitmix64(52, n)], axis=1).astype(np.uint64)
    raw = pairs.view(np.uint8).reshape(-1)
    offs = (np.arange(n + 1, dtype=np.uint64) * np.uint64(16))
    cfg = ph.make_config(lam, P, 0)
    r = ctx.build_bytes(raw, offs, cfg)
    d_p, d_out = hip_rt.DeviceBuffer.from_numpy(pairs), hip_rt.DeviceBuffer(n * 4)
    ctx.eval_pairs_device(d_p, n, cfg, r.np_, device_mphf(r, 1.0), d_out)
    hip_rt.synchronize()
    got = d_out.to_numpy(np.uint32)
    assert np.array_equal(got, oracle_positions(po.hash_bytes(raw, offs, 0), r, lam, P, 1.0))
    assert np.unique(got).size == n


def test_eval_prehashed_keys_equal_oracle(ctx):
    n, lam, P, alpha = 600_000, 7.5, 2048, 0.99
    hk = po.hash_u64(po.splitmix64(61, n), 5)
    cfg = ph.make_config(lam, P, 0, alpha)
    r = ctx.build_keys(hk, cfg)
    d_k, d_out = hip_rt.DeviceBuffer.from_numpy(hk), hip_rt.DeviceBuffer(n * 4)
    ctx.eval_keys_device(d_k, n, cfg, r.np_, device_mphf(r, alpha), d_out)
    hip_rt.synchronize()
    got = d_out.to_numpy(np.uint32)
    assert np.array_equal(got, oracle_positions(hk, r, lam, P, alpha))
    assert np.unique(got).size == n


def test_eval_key_in_an_empty_partition(ctx):
    """Pre-hashed keys placed in partitions 0 and 2 of 3 (P = 4): queries that
    fall into the empty partition 1 get offsets[1]; members keep their
    positions.  At alpha = 1 there is no remap table to read."""
    rng = np.random.default_rng(3)
    hk = rng.integers(0, 2**32, (9, 4), dtype=np.uint64).astype(np.uint32)
    hk[:5, 0] = rng.integers(0, 2**32 // 3 - 8, 5)            # partition 0 of 3
    hk[5:, 0] = rng.integers(2 * (2**32 // 3) + 8, 2**32, 4)  # partition 2 of 3
    cfg = ph.make_config(2.0, 4)
    r = ctx.build_keys(hk, cfg)
    assert list(np.diff(r.offsets.astype(np.int64))) == [5, 0, 4]
    q = rng.integers(0, 2**32, (64, 4), dtype=np.uint64).astype(np.uint32)
    q[:, 0] = rng.integers(2**32 // 3 + 8, 2 * (2**32 // 3) - 8, 64)  # partition 1
    allq = np.concatenate([hk, q])
    d_k, d_out = hip_rt.DeviceBuffer.from_numpy(allq), hip_rt.DeviceBuffer(allq.shape[0] * 4)
    ctx.eval_keys_device(d_k, allq.shape[0], cfg, r.np_, device_mphf(r, 1.0), d_out)
    hip_rt.synchronize()
    got = d_out.to_numpy(np.uint32)
    assert sorted(got[:9].tolist()) == list(range(9))
    assert np.all(got[9:] == r.offsets[1])
