"""The allocation-free entry point (phobic_build_u64_device_noalloc) against
phobic_build_u64_device, through the C ABI: bit-identical pilots, offsets and
remap tables on consecutive calls that share the fixed scratch array, the
second smaller and with another configuration; builds beyond the array are
refused before anything is enqueued."""
import numpy as np
import pytest

import hip_rt
import phobic_hip as ph

pytestmark = pytest.mark.gpu


def _outputs(cfg, n):
    np_, B = ph.sizes(cfg, n)
    rcap = ph.remap_capacity(cfg, n)
    bufs = [hip_rt.DeviceBuffer(max(np_ * B, 1) * 4), hip_rt.DeviceBuffer((np_ + 1) * 4),
            hip_rt.DeviceBuffer(max(rcap, 1) * 4) if rcap else None,
            hip_rt.DeviceBuffer((np_ + 1) * 4) if rcap else None]
    return bufs, ph.device_outputs(*bufs)


def _fetch(bufs, cfg, n):
    np_, B = ph.sizes(cfg, n)
    got = [bufs[0].to_numpy(np.uint32, np_ * B), bufs[1].to_numpy(np.uint32, np_ + 1)]
    if bufs[2] is not None:
        roff = bufs[3].to_numpy(np.uint32, np_ + 1)
        got += [roff, bufs[2].to_numpy(np.uint32, int(roff[-1]))]
    return got


def _check_pair(ctx, n, cfg, kseed, stream):
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(kseed, n, dk, stream)
    ref_bufs, ref_out = _outputs(cfg, n)
    na_bufs, na_out = _outputs(cfg, n)
    ds_ref, ds_na = hip_rt.DeviceBuffer(4), hip_rt.DeviceBuffer(4)
    ctx.build_u64_device(dk, n, cfg, ref_out, ds_ref, stream)
    ctx.build_u64_device_noalloc(dk, n, cfg, na_out, ds_na, stream)
    stream.synchronize()
    assert int(ds_ref.to_numpy(np.uint32)[0]) == 0
    assert int(ds_na.to_numpy(np.uint32)[0]) == 0
    for a, b in zip(_fetch(ref_bufs, cfg, n), _fetch(na_bufs, cfg, n)):
        assert np.array_equal(a, b)


def test_noalloc_equals_device_build_on_consecutive_calls(ctx):
    stream = hip_rt.Stream()
    print("noalloc call 1: 1.2M keys, lambda 7.5, P 2048, alpha 0.99", flush=True)
    _check_pair(ctx, 1_200_000, ph.make_config(7.5, 2048, 0, 0.99), 3, stream)
    print("noalloc call 2: 300k keys, lambda 7.0, P 2048, alpha 1", flush=True)
    _check_pair(ctx, 300_000, ph.make_config(7.0, 2048, 0, 1.0), 4, stream)
    print("noalloc call 3: 50k keys, lambda 8.0, P 2500, alpha 1", flush=True)
    _check_pair(ctx, 50_000, ph.make_config(8.0, 2500, 0, 1.0), 5, stream)


def test_noalloc_refuses_builds_beyond_its_scratch(ctx):
    n = 4_000_000  # about 24 * 3072 bytes per partition of 2048 keys: well above 64 MiB
    cfg = ph.make_config(7.0, 2048)
    bufs, out = _outputs(cfg, n)
    ds = hip_rt.DeviceBuffer(4)
    with pytest.raises(ph.PhobicError) as e:
        ctx.build_u64_device_noalloc(1 << 20, n, cfg, out, ds)  # refused before the keys are read
    assert e.value.status == 11
