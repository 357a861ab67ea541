"""Batched builds: every set's pilots and offsets equal building it alone
(and the oracle), eagerly and through replays of the captured hipGraph."""
import numpy as np
import pytest

import hip_rt
import phobic_hip as ph
import pyoracle as po

pytestmark = pytest.mark.gpu


def make_sets(count, lo, hi, seed, extra=()):
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    sizes = np.concatenate([sizes, np.array(extra, dtype=np.int64)])
    offs = np.zeros(sizes.size + 1, np.uint64)
    offs[1:] = np.cumsum(sizes)
    keys = np.concatenate([po.splitmix64(1000 + s, int(n)) for s, n in enumerate(sizes)] or [np.zeros(0, np.uint64)])
    return sizes, offs, keys


def run_batch(ctx, cfg, offs, keys, graph=False, reps=1, between=None):
    b = ph.Batch(ctx, cfg, offs)
    dk = hip_rt.DeviceBuffer.from_numpy(keys)
    dp = hip_rt.DeviceBuffer(max(b.total_pilots, 1) * 4)
    do = hip_rt.DeviceBuffer(b.total_offsets * 4)
    ds = hip_rt.DeviceBuffer(4)
    outs = []
    for r in range(reps):
        dp.zero()
        do.zero()
        b.launch(dk, dp, do, ds, graph=graph)
        hip_rt.synchronize()
        assert int(ds.to_numpy(np.uint32)[0]) == 0
        outs.append((dp.to_numpy(np.uint32, b.total_pilots), do.to_numpy(np.uint32, b.total_offsets)))
        if between is not None:
            between()
    return b, outs


def test_batch_sets_equal_individual_builds(ctx):
    cfg = ph.make_config(7.0, 2048)
    sizes, offs, keys = make_sets(120, 1e3, 2e5, 1, extra=(0, 1, 2048, 2049))
    b, [(pil, off)] = run_batch(ctx, cfg, offs, keys)
    for s in range(sizes.size):
        sk = keys[int(offs[s]):int(offs[s + 1])]
        r = ctx.build_u64(sk, cfg)
        bp, bo = b.split(pil, off, s)
        assert np.array_equal(bo, r.offsets), s
        assert np.array_equal(bp, r.pilots), s
    for s in (0, 7, sizes.size - 1):  # and the oracle for a few sets
        sk = keys[int(offs[s]):int(offs[s + 1])]
        st, opil, ooff = po.build_keys(po.hash_u64(sk, 0), 7.0, 2048)
        bp, bo = b.split(pil, off, s)
        assert np.array_equal(bp, opil) and np.array_equal(bo, ooff)
    b.close()


def test_batch_graph_replays_match_eager(ctx):
    cfg = ph.make_config(8.0, 2500)
    sizes, offs, keys = make_sets(60, 1e3, 1e5, 2)
    _, [eager] = run_batch(ctx, cfg, offs, keys)
    other = po.splitmix64(5, 300_000)

    def eager_work():  # unrelated build on the same context between replays
        ctx.build_u64(other, ph.make_config(7.0, 2048))

    b, outs = run_batch(ctx, cfg, offs, keys, graph=True, reps=4, between=eager_work)
    for pil, off in outs:
        assert np.array_equal(pil, eager[0]) and np.array_equal(off, eager[1])
    b.close()


def test_batch_single_set_equals_plain(ctx):
    cfg = ph.make_config(7.5, 2048)
    keys = po.splitmix64(3, 1_500_000)
    offs = np.array([0, keys.size], np.uint64)
    b, [(pil, off)] = run_batch(ctx, cfg, offs, keys)
    r = ctx.build_u64(keys, cfg)
    assert np.array_equal(pil, r.pilots) and np.array_equal(off, r.offsets)
    b.close()


def test_tuning_does_not_change_results(ctx):
    """Every launch shape gives the same pilots: one, two and three routing
    levels (1465 partitions), raw-key and hashed-record levels, tiles of 512
    and 1024 threads, and split tiles reduced to fit the LDS."""
    keys = po.splitmix64(21, 3_000_000)
    cfg = ph.make_config(7.5, 2048)
    saved = ctx.tuning()
    ref = ctx.build_u64(keys, cfg)
    try:
        shapes = [(1, 8, 32, 1, 0, 0), (2, 16, 64, 2, 0, 0), (8, 8, 128, 3, 0, 0), (4, 16, 16, 0, 0, 0),
                  (4, 16, 128, 3, 0, 0), (2, 12, 128, 0, 0, 0), (2, 12, 40, 1, 0, 0), (2, 16, 128, 0, 8, 32),
                  (2, 16, 64, 0, 16, 16), (1, 8, 128, 0, 12, 2), (2, 16, 16, 0, 8, 3),
                  (1, 20, 128, 0, 12, 48), (2, 20, 64, 0, 0, 32),
                  (1, 24, 1024, 0, 24, 0), (1, 24, 1024, 0, 20, 512), (1, 20, 1024, 0, 20, 48),
                  (1, 12, 1024, 0, 12, 0, 1024), (1, 8, 1024, 0, 4, 64, 1024), (1, 10, 40, 0, 6, 0, 1024),
                  (1, 24, 1024, 0, 24, 1024)]  # one level of 1024 bins: 1465 partitions need a second
        for shape in shapes:
            ctx.set_tuning(*shape)
            r = ctx.build_u64(keys, cfg)
            assert np.array_equal(r.pilots, ref.pilots) and np.array_equal(r.offsets, ref.offsets), shape
    finally:
        ctx.set_tuning(**saved)


def test_tuning_defaults_and_validation():
    """A new context starts with the documented shapes (include/phobic_hip.h);
    out-of-range shapes are refused and leave the tuning unchanged."""
    c = ph.Context(0)
    try:
        d = c.tuning()
        assert d == {"search_waves": 1, "route_per": 12, "max_fanout": 1024, "search_mode": 0, "split_per": 12,
                     "first_fanout": 0, "route_threads": 1024}
        for bad in ({"search_waves": 3}, {"route_per": 28}, {"route_per": 14}, {"split_per": 16},
                    {"max_fanout": 1025}, {"max_fanout": 8}, {"search_mode": 4}, {"first_fanout": 1},
                    {"first_fanout": 2000}, {"route_threads": 256}, {"route_per": 24}):
            t = dict(d)
            t.update(bad)
            with pytest.raises(ph.PhobicError):
                c.set_tuning(**t)
            assert c.tuning() == d, bad
        c.set_tuning(2, 16, 64, 1, 8, 32, 512)
        assert c.tuning()["route_per"] == 16 and c.tuning()["split_per"] == 8
    finally:
        c.close()


def test_autotune_returns_valid_shape(ctx):
    saved = ctx.tuning()
    best, ms = ctx.autotune(ph.make_config(7.0, 2048), 2_000_000)
    assert best["search_waves"] in (1, 2, 4, 8) and best["route_threads"] in (512, 1024)
    assert best["route_per"] in ((16, 20, 24) if best["route_threads"] == 512 else (8, 10, 12))
    assert best["max_fanout"] in (128, 1024) and best["search_mode"] in (0, 1, 2, 3) and ms > 0
    assert best["split_per"] in (8, 10, 12, 16, 20, 24) and best["first_fanout"] in (0, 48, 512)
    assert ctx.tuning() == best
    ctx.set_tuning(**saved)
