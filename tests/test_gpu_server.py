"""Serving: requests submitted concurrently while others run produce the same
results as building each one alone; admission respects the HBM budget."""
import threading

import numpy as np
import pytest

import phobic_hip as ph
import pyoracle as po

pytestmark = pytest.mark.gpu


def request_mix(count, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        lam = float(rng.choice([6.0, 7.0, 8.0]))
        alpha = float(rng.choice([0.99, 1.0]))
        cfg = ph.make_config(lam, 2048, 0, alpha)
        if rng.random() < 0.7:
            reqs.append(("u64", po.splitmix64(5000 + i, n), cfg))
        else:
            m = max(1000, n // 8)
            lens = rng.integers(8, 65, m)
            offs = np.zeros(m + 1, np.uint64)
            offs[1:] = np.cumsum(lens)
            data = rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8)
            # make strings distinct: overwrite 8 bytes with the index (printable hex)
            for j in range(m):
                data[int(offs[j]):int(offs[j]) + 8] = np.frombuffer(b"%08x" % j, np.uint8)
            reqs.append(("bytes", (data, offs), cfg))
    return reqs


def alone(ctx, kind, payload, cfg):
    if kind == "u64":
        return ctx.build_u64(payload, cfg)
    return ctx.build_bytes(payload[0], payload[1], cfg)


def same(a, b):
    ok = np.array_equal(a.pilots, b.pilots) and np.array_equal(a.offsets, b.offsets)
    if a.remap is not None or b.remap is not None:
        ok = ok and np.array_equal(a.remap, b.remap) and np.array_equal(a.remap_offsets, b.remap_offsets)
    return ok


def test_results_alone_equal_in_the_mix(ctx):
    reqs = request_mix(24, 1)
    srv = ph.Server(slots=4, hbm_budget=64 << 30)
    ids = [None] * len(reqs)

    def submit(lo, hi):  # two submitting threads, arrivals interleave with running builds
        for i in range(lo, hi):
            kind, payload, cfg = reqs[i]
            ids[i] = srv.submit_u64(payload, cfg) if kind == "u64" else srv.submit_bytes(payload[0], payload[1], cfg)

    th = [threading.Thread(target=submit, args=(0, 12)), threading.Thread(target=submit, args=(12, 24))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, (kind, payload, cfg) in enumerate(reqs):
        st, lat = srv.wait(ids[i])
        assert st == 0 and lat > 0
        mix = srv.fetch(ids[i])
        assert same(mix, alone(ctx, kind, payload, cfg)), i
        srv.release(ids[i])
    s = srv.stats()
    assert s["completed"] == len(reqs) and s["rejected"] == 0 and s["hbm_in_use"] == 0
    assert s["p99_ms"] >= s["p50_ms"] > 0
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
