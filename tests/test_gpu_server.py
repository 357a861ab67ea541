"""Serving: requests submitted while others run give the oracle's results,
alone or built together in one launch with other queued requests of their
key kind, partition size and load factor (lambda and the seed may differ
from request to request); admission keeps the server's counted memory within the budget
and never below what the device pool actually holds."""
import threading
import time

import numpy as np
import pytest

import hip_rt
import phobic_hip as ph
import pyoracle as po

pytestmark = pytest.mark.gpu


def request_mix(count, seed, lo=1e4, hi=2e6):
    """u64 (70 %) and printable string (30 %) requests, lambda {6,7,8},
    alpha {0.99, 1}, sizes log-uniform in [lo, hi] (strings: an eighth)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        lam = float(rng.choice([6.0, 7.0, 8.0]))
        alpha = float(rng.choice([0.99, 1.0]))
        cfg = ph.make_config(lam, 2048, 0, alpha)
        if rng.random() < 0.7:
            reqs.append(("u64", po.splitmix64(5000 + i, n), cfg, lam, alpha))
        else:
            m = max(1000, n // 8)
            offs = np.zeros(m + 1, np.uint64)
            offs[1:] = np.cumsum(rng.integers(8, 65, m))
            data = rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8)
            reqs.append(("bytes", (data, offs), cfg, lam, alpha))
    return reqs


def submit(srv, req):
    kind, payload, cfg = req[:3]
    return srv.submit_u64(payload, cfg) if kind == "u64" else srv.submit_bytes(payload[0], payload[1], cfg)


def check_against_oracle(r, req):
    """Pilots, offsets and (alpha < 1) the remap table equal the oracle's."""
    kind, payload, _, lam, alpha = req
    hk = po.hash_u64(payload, 0) if kind == "u64" else po.hash_bytes(payload[0], payload[1], 0)
    if alpha < 1.0:
        st, pil, off, rm, roff = po.build_keys_alpha(hk, lam, 2048, alpha)
        assert st == 0
        assert np.array_equal(r.remap, rm) and np.array_equal(r.remap_offsets, roff)
    else:
        st, pil, off = po.build_keys(hk, lam, 2048)
        assert st == 0
    assert np.array_equal(r.pilots, pil) and np.array_equal(r.offsets, off)


def test_results_alone_equal_in_the_mix():
    reqs = request_mix(24, 1)
    srv = ph.Server(slots=4, hbm_budget=64 << 30)
    ids = [None] * len(reqs)

    def go(lo, hi):  # two submitting threads: arrivals interleave with running launches
        for i in range(lo, hi):
            ids[i] = submit(srv, reqs[i])

    th = [threading.Thread(target=go, args=(0, 12)), threading.Thread(target=go, args=(12, 24))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, req in enumerate(reqs):
        st, lat = srv.wait(ids[i])
        assert st == 0 and lat > 0
        check_against_oracle(srv.fetch(ids[i]), req)
        srv.release(ids[i])
    s = srv.stats()
    assert s["completed"] == len(reqs) and s["rejected"] == 0
    assert s["p99_ms"] >= s["p50_ms"] > 0
    srv.close()


def test_queued_small_requests_share_launches():
    """One slot, busy with a large build while 60 small requests of three
    configurations queue behind it: they are built together, in launches of
    several requests each, and every result is the oracle's."""
    srv = ph.Server(slots=1, hbm_budget=64 << 30)
    big = srv.submit_u64(po.splitmix64(77, 30_000_000), ph.make_config(7.0, 2048))
    rng = np.random.default_rng(3)
    reqs = []
    for i in range(60):
        n = int(rng.integers(5_000, 150_000))
        kind = ["u64", "u64", "bytes"][i % 3]
        lam, alpha = [(7.0, 1.0), (7.5, 0.99), (6.0, 0.99)][i % 3]
        cfg = ph.make_config(lam, 2048, 0, alpha)
        if kind == "u64":
            reqs.append(("u64", po.splitmix64(9000 + i, n), cfg, lam, alpha))
        else:
            m = n // 4
            offs = np.zeros(m + 1, np.uint64)
            offs[1:] = np.cumsum(rng.integers(8, 65, m))
            data = rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8)
            reqs.append(("bytes", (data, offs), cfg, lam, alpha))
    ids = [submit(srv, r) for r in reqs]
    assert srv.wait(big)[0] == 0
    for rid, req in zip(ids, reqs):
        assert srv.wait(rid)[0] == 0
        check_against_oracle(srv.fetch(rid), req)
    s = srv.stats()
    assert s["grouped_launches"] >= 3 and s["grouped_requests"] >= 30
    for rid in ids + [big]:
        srv.release(rid)
    srv.close()


def test_large_requests_group_among_themselves():
    """Large requests (above 4e6 keys) of one configuration that queue behind
    a busy slot are built as one launch of their own, apart from the small
    requests of that configuration; host keys and device-resident keys mix
    in one group, and every result is the oracle's."""
    cfg = ph.make_config(8.0, 2048, 0, 0.99)
    sizes = [4_500_001, 6_100_000, 5_000_000, 120_000, 90_000, 4_000_000]
    keys = [po.splitmix64(400 + i, n) for i, n in enumerate(sizes)]
    d_keys = hip_rt.DeviceBuffer.from_numpy(keys[1])  # device-resident, between two host requests
    big = po.splitmix64(78, 60_000_000)
    srv = ph.Server(slots=1, hbm_budget=64 << 30)
    blocker = srv.submit_u64(big, ph.make_config(7.0, 2048))  # the others queue behind it
    ids = [srv.submit_u64(d_keys, cfg, sizes[1]) if i == 1 else srv.submit_u64(k, cfg) for i, k in enumerate(keys)]
    assert srv.wait(blocker)[0] == 0
    for rid, k in zip(ids, keys):
        assert srv.wait(rid)[0] == 0
        check_against_oracle(srv.fetch(rid), ("u64", k, cfg, 8.0, 0.99))
    s = srv.stats()
    # one launch of the three large requests, one of the three small ones
    assert s["grouped_launches"] == 2 and s["grouped_requests"] == 6
    for rid in ids + [blocker]:
        srv.release(rid)
    srv.close()


@pytest.mark.parametrize("kind", ["u64", "bytes"])
def test_mixed_lambdas_and_seeds_share_a_launch(kind):
    """Small requests of one partition size and load factor but five lambdas
    (four of each) and several seeds, queued behind a busy slot: launches of
    up to four lambdas build them together (at least sixteen requests in
    groups, whichever request the slot starts alone if the blocker ends
    before the rest are queued), and every result equals the oracle's for
    the request's own lambda and seed."""
    srv = ph.Server(slots=1, hbm_budget=64 << 30)
    blocker = srv.submit_u64(po.splitmix64(80, 60_000_000), ph.make_config(7.0, 2048))
    rng = np.random.default_rng(5)
    lams = [6.0, 7.0, 7.5, 8.0, 9.0]
    reqs = []
    for i in range(20):
        n = int(rng.integers(5_000, 200_000))
        lam = lams[i % 5]
        seed = int(rng.integers(1, 1 << 63)) if i % 2 else 0
        cfg = ph.make_config(lam, 2048, seed, 0.99)
        if kind == "u64":
            reqs.append((po.splitmix64(300 + i, n), None, cfg, lam, seed))
        else:
            m = n // 4
            offs = np.zeros(m + 1, np.uint64)
            offs[1:] = np.cumsum(rng.integers(8, 65, m))
            reqs.append((rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8), offs, cfg, lam, seed))
    ids = [srv.submit_u64(a, cfg) if kind == "u64" else srv.submit_bytes(a, b, cfg) for a, b, cfg, _, _ in reqs]
    assert srv.wait(blocker)[0] == 0
    for rid, (a, b, _, lam, seed) in zip(ids, reqs):
        assert srv.wait(rid)[0] == 0
        hk = po.hash_u64(a, seed) if kind == "u64" else po.hash_bytes(a, b, seed)
        st, pil, off, rm, roff = po.build_keys_alpha(hk, lam, 2048, 0.99)
        assert st == 0
        r = srv.fetch(rid)
        assert np.array_equal(r.pilots, pil) and np.array_equal(r.offsets, off)
        assert np.array_equal(r.remap, rm) and np.array_equal(r.remap_offsets, roff)
    s = srv.stats()
    assert 2 <= s["grouped_launches"] <= 3 and s["grouped_requests"] >= 16
    for rid in ids + [blocker]:
        srv.release(rid)
    srv.close()


def test_device_resident_sets_read_in_place():
    """A group of u64 requests, some in device memory (read where they lie)
    and some in host memory (staged), with sets of 1 to 200,000 keys, so
    that routing tiles span many sets: every result is the oracle's."""
    rng = np.random.default_rng(21)
    sizes = [1, 3, 2, 100, 5000, 1, 70_000, 17, 200_000, 9, 4096, 8191, 8193, 1, 150_000, 33]
    srv = ph.Server(slots=1, hbm_budget=64 << 30)
    blocker = srv.submit_u64(po.splitmix64(81, 60_000_000), ph.make_config(7.0, 2048))
    reqs, bufs, ids = [], [], []
    for i, n in enumerate(sizes):
        lam = [6.0, 7.5, 8.0][i % 3]
        seed = int(rng.integers(0, 1 << 63))
        cfg = ph.make_config(lam, 2048, seed, 1.0)
        keys = po.splitmix64(600 + i, n)
        if i % 2:
            d = hip_rt.DeviceBuffer.from_numpy(keys)
            bufs.append(d)
            ids.append(srv.submit_u64(d, cfg, n))
        else:
            ids.append(srv.submit_u64(keys, cfg))
        reqs.append((keys, lam, seed))
    assert srv.wait(blocker)[0] == 0
    for rid, (keys, lam, seed) in zip(ids, reqs):
        assert srv.wait(rid)[0] == 0
        st, pil, off = po.build_keys(po.hash_u64(keys, seed), lam, 2048)
        assert st == 0
        r = srv.fetch(rid)
        assert np.array_equal(r.pilots, pil) and np.array_equal(r.offsets, off)
    assert srv.stats()["grouped_requests"] >= 12
    for rid in ids + [blocker]:
        srv.release(rid)
    srv.close()


def test_grouped_host_strings_rebased_through_staging():
    """A group of host string requests larger than the direct-copy limit
    (their bytes and offsets go through the pinned staging chunks, the
    offsets rebased per request while staged), with requests whose offsets
    do not start at 0, gives the oracle's results."""
    cfg = ph.make_config(6.0, 2048, 0, 0.99)
    rng = np.random.default_rng(11)
    reqs = []
    for n in [600_000, 1_500_000, 35_000, 900_000]:
        skip = int(rng.integers(0, 5000))  # bytes before the first key
        offs = np.zeros(n + 1, np.uint64)
        offs[1:] = np.cumsum(rng.integers(8, 65, n))
        offs += skip
        data = rng.integers(0x20, 0x7F, int(offs[-1]) + int(rng.integers(0, 100)), dtype=np.uint8)
        reqs.append(("bytes", (data, offs), cfg, 6.0, 0.99))
    big = po.splitmix64(79, 60_000_000)
    srv = ph.Server(slots=1, hbm_budget=64 << 30)
    blocker = srv.submit_u64(big, ph.make_config(7.0, 2048))  # the others queue behind it
    ids = [srv.submit_bytes(r[1][0], r[1][1], cfg) for r in reqs]
    assert srv.wait(blocker)[0] == 0
    for rid, req in zip(ids, reqs):
        assert srv.wait(rid)[0] == 0
        check_against_oracle(srv.fetch(rid), req)
    s = srv.stats()
    assert s["grouped_launches"] == 1 and s["grouped_requests"] == 4
    for rid in ids + [blocker]:
        srv.release(rid)
    srv.close()


def test_counted_memory_covers_the_pool_during_a_trace():
    """Sampled while a mixed trace runs: the server's counted device memory is
    never below what the library's pool holds beyond its level before the
    server existed, and never above the budget."""
    budget = 12 << 30
    reqs = request_mix(40, 2, 1e4, 8e6)
    srv = ph.Server(slots=4, hbm_budget=budget)
    base = srv.stats()["pool_used"]
    samples, stop = [], threading.Event()

    def sample():
        while not stop.is_set():
            s = srv.stats()
            samples.append((s["hbm_in_use"], s["pool_used"]))
            time.sleep(0.0002)

    t = threading.Thread(target=sample)
    t.start()
    try:
        ids = [submit(srv, r) for r in reqs]
        for i, rid in enumerate(ids):
            assert srv.wait(rid)[0] == 0
            if i % 2:
                srv.release(rid)
        for i, rid in enumerate(ids):
            if i % 2 == 0:
                srv.release(rid)
    finally:
        stop.set()
        t.join()
    assert len(samples) >= 10  # taken while the trace ran
    for counted, used in samples:
        assert counted >= used - base, (counted, used, base)
        assert counted <= budget
    srv.close()


def test_budget_admission_and_rejection():
    srv = ph.Server(slots=2, hbm_budget=256 << 20)  # 256 MiB
    keys = po.splitmix64(7, 200_000)
    rid = srv.submit_u64(keys, ph.make_config(7.0, 2048))
    assert srv.wait(rid)[0] == 0
    srv.release(rid)
    with pytest.raises(ph.PhobicError) as e:  # alone larger than the budget
        srv.submit_u64(po.splitmix64(8, 20_000_000), ph.make_config(7.0, 2048))
    assert e.value.status == 3
    assert srv.stats()["rejected"] == 1
    srv.close()


def test_malformed_string_requests_refused_at_submit():
    """Host string requests whose offsets decrease, or end past the bytes
    given, are refused before anything reaches the device."""
    srv = ph.Server(slots=1, hbm_budget=1 << 30)
    cfg = ph.make_config(6.0, 2048)
    data = np.frombuffer(b"abcdefghijklmnopqrstuvwxyz012345", np.uint8)
    with pytest.raises(ph.PhobicError) as e:
        srv.submit_bytes(data, np.array([0, 8, 4, 16], np.uint64), cfg)
    assert e.value.status == 1
    with pytest.raises(ph.PhobicError) as e:
        srv.submit_bytes(data, np.array([0, 8, 16, 40], np.uint64), cfg)
    assert e.value.status == 1
    rid = srv.submit_bytes(data, np.array([0, 8, 16, 32], np.uint64), cfg)
    assert srv.wait(rid)[0] == 0
    srv.release(rid)
    srv.close()


def test_failed_request_reports_status_and_server_continues():
    srv = ph.Server(slots=2, hbm_budget=8 << 30)
    bad = po.splitmix64(9, 50_000)
    bad[100] = bad[7]
    rb = srv.submit_u64(bad, ph.make_config(7.0, 2048))
    good = po.splitmix64(10, 80_000)
    rg = srv.submit_u64(good, ph.make_config(7.0, 2048))
    assert srv.wait(rb)[0] == 7
    assert srv.wait(rg)[0] == 0
    hk = po.hash_u64(good, 0)
    st, pil, off = po.build_keys(hk, 7.0, 2048)
    r = srv.fetch(rg)
    assert np.array_equal(r.pilots, pil) and np.array_equal(r.offsets, off)
    srv.release(rb)
    srv.release(rg)
    srv.close()
