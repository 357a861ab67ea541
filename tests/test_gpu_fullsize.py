"""Bit-exact parity at the full sizes of the headline workloads, through the
C ABI on an MI355X.

* 100M u64 keys (splitmix64 seed 2), lambda 8, P 2500, alpha 1: every pilot
  and every partition offset equals the oracle's.
* 1e9 u64 keys (splitmix64 seed 3), lambda 7.5, P 2048, alpha 0.99: every
  partition offset and remap offset equals the oracle's scan of the full key
  stream; pilots and remap entries are compared on a sample of partitions
  rebuilt by the oracle from the full key stream: 1% at random, every
  partition whose largest bucket holds 150 keys or more, every partition
  with a pilot index of 1024 or more in the device build, and the first and
  last partitions.  A partition depends only on its own keys, so each
  sampled row is the full build's row.

The oracle streams the keys (it never holds the billion keys), so these run
in well under a minute of CPU time on the test host.
"""
import numpy as np
import pytest

import hip_rt
import phobic_hip as ph
import pyoracle as po

pytestmark = pytest.mark.gpu


def device_build(ctx, kseed, n, cfg):
    np_, B = ph.sizes(cfg, n)
    rcap = ph.remap_capacity(cfg, n)
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(kseed, n, dk)
    dp, do, ds = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    drm = hip_rt.DeviceBuffer(max(rcap, 1) * 4) if rcap else None
    dro = hip_rt.DeviceBuffer((np_ + 1) * 4) if rcap else None
    ctx.build_u64_device(dk, n, cfg, ph.device_outputs(dp, do, drm, dro), ds)
    hip_rt.synchronize()
    assert int(ds.to_numpy(np.uint32)[0]) == 0
    del dk
    out = {"pilots": dp.to_numpy(np.uint32), "offsets": do.to_numpy(np.uint32)}
    if rcap:
        roff = dro.to_numpy(np.uint32)
        out["remap_offsets"] = roff
        out["remap"] = drm.to_numpy(np.uint32, int(roff[-1]))
    return out


def table_sizes(counts, alpha):
    a = po.alpha_ppm(alpha)
    c = counts.astype(np.int64)
    return c if a >= 1_000_000 else (c * 1_000_000 + a - 1) // a


@pytest.mark.timeout(900)
def test_100m_every_pilot_and_offset(ctx):
    n, lam, P, kseed = 100_000_000, 8.0, 2500, 2
    cfg = ph.make_config(lam, P)
    np_, B = ph.sizes(cfg, n)
    print(f"100M keys, lambda {lam}, P {P}: device build", flush=True)
    got = device_build(ctx, kseed, n, cfg)
    print("oracle: every partition", flush=True)
    counts, _ = po.scan_splitmix(kseed, n, 0, lam, P, max_bucket=False)
    off = np.zeros(np_ + 1, np.int64)
    off[1:] = np.cumsum(counts)
    assert np.array_equal(got["offsets"].astype(np.int64), off)
    st, pil, cnt, _ = po.build_splitmix_parts(kseed, n, 0, lam, P, 1.0, np.arange(np_, dtype=np.uint32))
    assert st == 0 and np.array_equal(cnt, counts)
    assert np.array_equal(got["pilots"].reshape(B, np_), pil.T)


@pytest.mark.timeout(900)
def test_1b_alpha099_offsets_and_sampled_partitions(ctx):
    n, lam, P, alpha, kseed = 1_000_000_000, 7.5, 2048, 0.99, 3
    cfg = ph.make_config(lam, P, 0, alpha)
    np_, B = ph.sizes(cfg, n)
    print(f"1e9 keys, lambda {lam}, P {P}, alpha {alpha}: device build", flush=True)
    got = device_build(ctx, kseed, n, cfg)
    print("oracle: scan of the full key stream", flush=True)
    counts, maxb = po.scan_splitmix(kseed, n, 0, lam, P)
    off = np.zeros(np_ + 1, np.int64)
    off[1:] = np.cumsum(counts)
    assert np.array_equal(got["offsets"].astype(np.int64), off)
    m = table_sizes(counts, alpha)
    roff = np.zeros(np_ + 1, np.int64)
    roff[1:] = np.cumsum(m - counts)
    assert np.array_equal(got["remap_offsets"].astype(np.int64), roff)

    pil = got["pilots"].reshape(B, np_)  # pil[b, p]
    pilot_index = (pil.astype(np.int64) // np.maximum(m, 1)[None, :]).max(axis=0)
    rng = np.random.default_rng(7)
    sel = set(rng.choice(np_, np_ // 100, replace=False).tolist())
    big = np.nonzero(maxb >= 150)[0]
    far = np.nonzero(pilot_index >= 1024)[0]
    sel.update(big.tolist())
    sel.update(far.tolist())
    sel.update([0, 1, np_ - 2, np_ - 1])
    parts = np.array(sorted(sel), np.uint32)
    print(f"oracle: {parts.size} partitions ({big.size} with a bucket of >= 150 keys, {far.size} with a pilot "
          f"index >= 1024, largest bucket {int(maxb.max())}, table sizes {int(m.min())}..{int(m.max())})",
          flush=True)
    assert far.size > 0  # the general pilot loop is exercised at this size
    st, opil, ocnt, orm = po.build_splitmix_parts(kseed, n, 0, lam, P, alpha, parts)
    assert st == 0
    assert np.array_equal(ocnt, counts[parts])
    assert np.array_equal(pil[:, parts].T, opil)
    rm = got["remap"]
    want = np.concatenate([rm[roff[p]:roff[p + 1]] for p in parts.tolist()])
    assert np.array_equal(want, orm)
