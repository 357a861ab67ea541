"""Benchmark: PHOBIC construction on one MI355X.

One step = one full device build (hash -> partition -> bucket sort ->
pilot search -> slot remapping) of n distinct synthetic u64 keys that are
already resident in HBM, producing pilots u32[np*B], partition offsets
u32[np+1] and, for alpha < 1, the remap table in HBM.  The headline workload
is 1e9 keys, lambda 7.5, partition size 2048, alpha 0.99.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode all|plain|batch|serve|strings|copy]

The default run (--mode all) prints ONE JSON line: the headline build
(Mkeys/s, per-stage device times, the streaming stages against their HBM
roofline, the search kernel against its instruction-issue roofline, device
query throughput, a CPU baseline of the oracle on a bounded sample), plus
the other figures of the metric: `batched_vs_plain` (4096 uneven key sets
built in one batched launch against one synchronous build per set),
`serving` (a Poisson request trace through the build server: p50 / p99
latency and served against sequential Mkeys/s) and `strings_200m` (the
string-keyed build).  The other modes print one line for their part alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "phobic-gpu_amd"))

import numpy as np  # noqa: E402

import hip_rt  # noqa: E402
import phobic_hip as ph  # noqa: E402

HBM_PEAK_GBS = 8000.0
# wave64 VALU instructions per second: 256 CUs x 4 SIMDs, one instruction
# per SIMD every 4 cycles (16 lanes per cycle) at the 2.4 GHz peak clock
VALU_PEAK_GINST = 256 * 4 * 2.4e9 / 4 / 1e9


def stage_bytes_per_key(name: str, lam: float, P: int, key_bytes: float = 8.0) -> float:
    """Algorithmic bytes per key of each kernel: the rows of the reference's
    stage structure that kernel performs (per-key figures: S1 k+1+4/l,
    S2 12/l+1028/P, S3 8/l+8/P, S4 8+1+4/l+8, S5 8+1028/P+12/l+4/l, where k
    is the key's input bytes: 8 for u64 keys, the mean string length + its
    8-byte offset for strings; from S1 on every key is its 64-bit hash, so
    the redistribution moves 8 bytes per key whatever the input was)."""
    s1 = key_bytes + 1 + 4 / lam
    s2 = 12 / lam + 1028 / P
    s3_apply, s3_scan = 8 / lam, 8 / P
    s4 = 8 + 1 + 4 / lam + 8
    s5 = 8 + 1028 / P + 16 / lam
    return {
        "hash_route": s1,               # hash + map; level-1 routing
        "partition_split": s4,          # routing to partitions (the redistribution)
        "partition_offsets": s3_scan,
        "partition_search": s2 + s3_apply + s5,
    }.get(name, 0.0)


def streaming_bytes_per_key(lam: float, P: int, key_bytes: float = 8.0) -> float:
    """Canonical S1-S4 bytes per key: 26 + 28/lambda + 1036/P for u64 keys
    (k + 18 + 28/lambda + 1036/P for k input bytes per key)."""
    return key_bytes + 18 + 28 / lam + 1036 / P


def profiled(kind: str, n: int, lam: float, P: int, alpha: float, **match) -> dict | None:
    """Figures that rocprofv3 runs recorded in profiles/*.md for this
    workload: lines `<kind> key=value ...` (HBM traffic from FETCH_SIZE /
    WRITE_SIZE passes scaled by a copy kernel of known size, or the search
    kernel's issue counters); None when no profile covers it."""
    import glob

    want = {"n": str(n), "lambda": str(lam), "P": str(P), "alpha": str(float(alpha))}
    want.update({k: str(v) for k, v in match.items()})
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if not line.startswith(kind + " "):
                continue
            kv = dict(tok.split("=", 1) for tok in line.split()[1:])
            if all(kv.get(k) == v for k, v in want.items()):
                return kv
    return None


def stage_report(kernels: dict, n: int, lam: float, P: int, alpha: float, key_bytes: float = 8.0) -> dict:
    """Per stage: device ms, Mkeys/s, algorithmic bytes per key and GB/s with
    their fraction of the 8 TB/s HBM peak, and the measured HBM GB/s where a
    profile of this workload recorded the stage's traffic."""
    rep = {}
    for k, ms in kernels.items():
        bpk = stage_bytes_per_key(k, lam, P, key_bytes)
        gbs = bpk * n / (ms / 1e3) / 1e9
        e = {"ms": round(ms, 4), "mkeys_s": round(n / (ms / 1e3) / 1e6, 1), "algo_bytes_per_key": round(bpk, 3),
             "algo_gbs": round(gbs, 1), "algo_hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        tr = profiled("traffic_bytes_per_stage", n, lam, P, alpha, stage=k)
        if tr:
            mg = float(tr["bytes"]) / (ms / 1e3) / 1e9
            e["measured_bytes_per_key"] = round(float(tr["bytes"]) / n, 2)
            e["measured_gbs"] = round(mg, 1)
            e["measured_hbm_frac"] = round(mg / HBM_PEAK_GBS, 4)
        rep[k] = e
    return rep


def cpu_baseline(n_sample: int, lam: float, P: int, alpha: float, kseed: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    keys = po.splitmix64(kseed, n_sample)
    t0 = time.perf_counter()
    hk = po.hash_u64(keys, 0)
    if alpha < 1.0:
        st = po.build_keys_alpha(hk, lam, P, alpha)[0]
    else:
        st = po.build_keys(hk, lam, P)[0]
    dt = time.perf_counter() - t0
    if st != 0:
        raise RuntimeError(f"oracle baseline failed with status {st}")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": round(n_sample / dt / 1e6, 3),
        "unit": "Mkeys/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_sample} keys (splitmix64 seed {kseed}) of the same configuration: oracle XXH3 hash + "
                  f"build (bucketing, bucket sort, pilot search{', slot remapping' if alpha < 1.0 else ''}), "
                  f"OpenMP over partitions, {dt:.1f} s",
    }


def apply_tuning_flags(ctx, args) -> None:
    """The --search-mode / --search-waves / ... overrides, on top of ctx's tuning."""
    overrides = {"search_mode": args.search_mode, "search_waves": args.search_waves, "route_per": args.route_per,
                 "max_fanout": args.max_fanout, "split_per": args.split_per, "first_fanout": args.first_fanout,
                 "route_threads": args.route_threads}
    if any(v is not None for v in overrides.values()):
        t = ctx.tuning()
        t.update({k: v for k, v in overrides.items() if v is not None})
        ctx.set_tuning(**t)


def run_plain(args) -> dict:
    """The headline: K timed device builds of 1e9 u64 keys resident in HBM."""
    n, lam, P, alpha = args.n, args.lam, args.P, args.alpha
    cfg = ph.make_config(lam, P, 0, alpha)
    np_, B = ph.sizes(cfg, n)
    rcap = ph.remap_capacity(cfg, n)
    ctx = ph.Context(0)
    tuning = None
    if args.autotune:
        tuning, _ = ctx.autotune(cfg, args.autotune)
    apply_tuning_flags(ctx, args)
    stream = hip_rt.Stream()
    dk = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(args.key_seed, n, dk, stream)
    dp = hip_rt.DeviceBuffer(np_ * B * 4)
    do = hip_rt.DeviceBuffer((np_ + 1) * 4)
    ds = hip_rt.DeviceBuffer(4)
    drm = hip_rt.DeviceBuffer(rcap * 4) if rcap else None
    dro = hip_rt.DeviceBuffer((np_ + 1) * 4) if rcap else None
    mphf = ph.device_outputs(dp, do, drm, dro)
    stream.synchronize()

    for _ in range(args.warmup):
        ctx.build_u64_device(dk, n, cfg, mphf, ds, stream)
    stream.synchronize()
    status = int(ds.to_numpy(np.uint32)[0])
    if status:
        raise RuntimeError(f"build failed: {ph.STATUS_NAMES.get(ph.status_from_bits(status))}")

    stage_sum: dict[str, float] = {}
    e0, e1 = hip_rt.Event(), hip_rt.Event()
    stream.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        ctx.build_u64_device(dk, n, cfg, mphf, ds, stream)
        for k, v in ctx.last_stage_times().items():  # waits for this build's events
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_ms(e1)
    status = int(ds.to_numpy(np.uint32)[0])
    if status:
        raise RuntimeError(f"build failed: {ph.STATUS_NAMES.get(ph.status_from_bits(status))}")

    ms_per_step = ev_ms / args.steps
    stages = {k: v / args.steps for k, v in stage_sum.items()}
    kernels = {k: v for k, v in stages.items() if k != "total"}
    # the streaming stages (S1-S4 of the reference's structure) against the
    # canonical bytes they must move
    stream_ms = kernels["hash_route"] + kernels["partition_split"] + kernels["partition_offsets"]
    sbpk = streaming_bytes_per_key(lam, P)
    s_gbs = sbpk * n / (stream_ms / 1e3) / 1e9
    streaming = {"stages": ["hash_route", "partition_split", "partition_offsets"], "ms": round(stream_ms, 4),
                 "algo_bytes_per_key": round(sbpk, 3), "achieved_gbs": round(s_gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                 "frac": round(s_gbs / HBM_PEAK_GBS, 4)}
    mtr = [profiled("traffic_bytes_per_stage", n, lam, P, alpha, stage=k) for k in streaming["stages"][:2]]
    if all(mtr):
        mb = sum(float(t["bytes"]) for t in mtr)
        streaming["measured_bytes_per_key"] = round(mb / n, 2)
        streaming["measured_gbs"] = round(mb / (stream_ms / 1e3) / 1e9, 1)
    # the search kernel is bound by instruction issue: VALU wave-instructions
    # per launch (counted by rocprofv3 for this workload) over its live
    # device time, against the VALU issue peak; its HBM figures beside
    search_ms = kernels["partition_search"]
    bpk = stage_bytes_per_key("partition_search", lam, P)
    hbm_gbs = bpk * n / (search_ms / 1e3) / 1e9
    roofline = {"kernel": "partition_search", "bound": "valu_issue"}
    issue = profiled("search_issue", n, lam, P, alpha, kernel="partition_search")
    if issue:
        g = float(issue["valu_insts"]) / (search_ms / 1e3) / 1e9
        roofline.update({"achieved": round(g, 1), "peak": VALU_PEAK_GINST, "unit": "G wave64 VALU inst/s",
                         "frac": round(g / VALU_PEAK_GINST, 4),
                         "valu_busy": float(issue["valu_busy"]) / 100, "salu_busy": float(issue["salu_busy"]) / 100,
                         "lds_util": float(issue["lds_util"]) / 100})
    tr = profiled("traffic_bytes_per_stage", n, lam, P, alpha, stage="partition_search")
    roofline["traffic"] = float(tr["bytes"]) if tr else None
    roofline["hbm"] = {"achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(hbm_gbs / HBM_PEAK_GBS, 4), "algo_bytes_per_key": round(bpk, 3)}
    out = {
        "metric": "MPHF build Mkeys/s at 1B uint64 keys",
        "value": round(n / (ms_per_step / 1e3) / 1e6, 2),
        "unit": "Mkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": f"synthetic: {n} distinct u64 keys, splitmix64 seed {args.key_seed}, generated in HBM",
        "config": {
            "workload": f"plain PHOBIC build of {n} u64 keys, lambda {lam}, partition size {P}, alpha {alpha}, "
                        "XXH3 seed 0, inputs and outputs resident in HBM",
            "n": n, "lambda": lam, "partition_size": P, "alpha": alpha, "partitions": np_,
            "buckets_per_partition": B,
        },
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "stages": stage_report(kernels, n, lam, P, alpha),
        "streaming": streaming,
        "tuning": ctx.tuning(),
        "autotuned": bool(tuning),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 3),
        "roofline": roofline,
    }
    if not args.no_verify:
        # every key queried on the device (MPHF::operator(), the reference's
        # query_time, benchmark.cpp:64-78) and the positions checked as a
        # bijection onto [0, n)
        out_pos = hip_rt.DeviceBuffer(n * 4)
        ctx.eval_u64_device(dk, n, cfg, np_, mphf, out_pos, stream)  # warm-up
        q0, q1 = hip_rt.Event(), hip_rt.Event()
        q0.record(stream)
        ctx.eval_u64_device(dk, n, cfg, np_, mphf, out_pos, stream)
        q1.record(stream)
        bm, bad = hip_rt.DeviceBuffer((n + 31) // 32 * 4), hip_rt.DeviceBuffer(8)
        ctx.check_bijection_device(out_pos, n, n, bm, bad, stream)
        stream.synchronize()
        qms = q0.elapsed_ms(q1)
        out["query"] = {"keys": n, "ms": round(qms, 3), "mkeys_s": round(n / (qms / 1e3) / 1e6, 1),
                        "ns_per_key": round(qms * 1e6 / n, 4), "order": "key order (random partitions and buckets)"}
        out["bijection_violations"] = int(bad.to_numpy(np.uint64)[0])
        del out_pos, bm, bad
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, lam, P, alpha, args.key_seed)
    del dk, dp, do, drm, dro, mphf
    ctx.close()
    return out


def run_batch(args) -> dict:
    """Batched builds of `sets` independent key sets (sizes log-uniform in
    [1e3, 1e6], seeded; keys of set s: splitmix64 seed 1000 + s), lambda 7,
    P 2048, in one launch (eager and as a hipGraph replay), against building
    the same sets one at a time with a synchronous device build each."""
    rng = np.random.default_rng(4)
    sizes = np.floor(np.exp(rng.uniform(np.log(1e3), np.log(1e6), args.sets))).astype(np.int64)
    offs = np.zeros(args.sets + 1, np.uint64)
    offs[1:] = np.cumsum(sizes)
    n = int(offs[-1])
    cfg = ph.make_config(7.0, 2048, 0)
    ctx = ph.Context(0)
    tuning = None
    if args.autotune:
        tuning, _ = ctx.autotune(cfg, args.autotune)
    stream = hip_rt.Stream()
    dk = hip_rt.DeviceBuffer(n * 8)
    for s in range(args.sets):
        ph.gen_splitmix64_device(1000 + s, int(sizes[s]), dk.ptr + 8 * int(offs[s]), stream)
    batch = ph.Batch(ctx, cfg, offs)
    dp_b = hip_rt.DeviceBuffer(batch.total_pilots * 4)
    do_b = hip_rt.DeviceBuffer(batch.total_offsets * 4)
    dp_p = hip_rt.DeviceBuffer(batch.total_pilots * 4)
    do_p = hip_rt.DeviceBuffer(batch.total_offsets * 4)
    ds = hip_rt.DeviceBuffer(4)
    stream.synchronize()

    def plain_once():
        for s in range(args.sets):
            pb = int(batch.part_base[s])
            out = ph.device_outputs(dp_p.ptr + 4 * batch.B * pb, do_p.ptr + 4 * (pb + s))
            ctx.build_u64_device(dk.ptr + 8 * int(offs[s]), int(sizes[s]), cfg, out, ds, stream)
            stream.synchronize()  # one synchronous build per set, as MPHFbuilder::build

    def timed(fn, reps):
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        stream.synchronize()
        return (time.perf_counter() - t0) / reps

    plain_once()  # warm-up (allocations)
    t_plain = timed(plain_once, 1)
    batch.launch(dk, dp_b, do_b, ds, stream)  # warm-up
    t_eager = timed(lambda: batch.launch(dk, dp_b, do_b, ds, stream), max(1, args.steps))
    batch.launch(dk, dp_b, do_b, ds, stream, graph=True)  # capture
    t_graph = timed(lambda: batch.launch(dk, dp_b, do_b, ds, stream, graph=True), max(1, args.steps))
    status = int(ds.to_numpy(np.uint32)[0])
    same = bool(np.array_equal(dp_b.to_numpy(np.uint32), dp_p.to_numpy(np.uint32)) and
                np.array_equal(do_b.to_numpy(np.uint32), do_p.to_numpy(np.uint32)))
    best = min(t_eager, t_graph)
    res = {
        "workload": f"{args.sets} independent u64 key sets, sizes floor(exp(U[ln 1e3, ln 1e6])) (numpy seed 4), "
                    "keys of set s = splitmix64 seed 1000+s in HBM, lambda 7.0, P 2048, alpha 1.0",
        "keys": n, "sets": args.sets,
        "value": round(n / best / 1e6, 2), "unit": "Mkeys/s",
        "plain_one_at_a_time_mkeys_s": round(n / t_plain / 1e6, 2),
        "batched_eager_mkeys_s": round(n / t_eager / 1e6, 2),
        "batched_graph_mkeys_s": round(n / t_graph / 1e6, 2),
        "speedup_vs_plain": round(t_plain / best, 2),
        "identical_outputs": same,
        "status": status,
        "tuning": tuning or ctx.tuning(),
        "autotuned": bool(tuning),
    }
    batch.close()
    del dk, dp_b, do_b, dp_p, do_p
    ctx.close()
    return res


def serve_trace(R: int, seed: int = 5):
    """The request trace: Poisson arrivals (mean 2000 requests/s), sizes
    log-uniform in [1e4, 1e8] (u64) / [1e4, 2e6] (strings of length 8..64,
    printable random bytes), 70 % u64 / 30 % strings, lambda in {6,7,8},
    alpha in {0.99, 1}; u64 keys generated in HBM, strings in host memory."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, R))
    reqs = []
    for i in range(R):
        lam = float(rng.choice([6.0, 7.0, 8.0]))
        alpha = float(rng.choice([0.99, 1.0]))
        cfg = ph.make_config(lam, 2048, 0, alpha)
        if rng.random() < 0.7:
            n = int(np.exp(rng.uniform(np.log(1e4), np.log(1e8))))
            d = hip_rt.DeviceBuffer(n * 8)
            ph.gen_splitmix64_device(20000 + i, n, d)
            reqs.append(("u64", n, d, cfg))
        else:
            n = int(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
            offs = np.zeros(n + 1, np.uint64)
            offs[1:] = np.cumsum(rng.integers(8, 65, n))
            data = rng.integers(0x20, 0x7F, int(offs[-1]), dtype=np.uint8)
            reqs.append(("bytes", n, (data, offs), cfg))
    hip_rt.synchronize()
    return arrivals, reqs


def run_serve(args) -> dict:
    """A finite Poisson trace of build requests served under a 200 GB HBM
    budget, against the plain sequential path (one synchronous device build
    at a time, in arrival order, over the same arrivals; host strings
    uploaded first, as the server uploads them).  Served throughput
    cannot exceed the trace's offered load (keys over the arrival span);
    the same requests submitted at once (`burst`) measure the server's
    capacity against the sum of the sequential build times.  The sequential
    path runs the trace twice and keeps each request's faster time; the
    server runs it three times each way after a warm-up pass, and the pass
    of median span is reported."""
    R = args.requests
    arrivals, reqs = serve_trace(R)
    total = sum(r[1] for r in reqs)

    # The sequential path serves each request with what the server does for
    # it: host strings are uploaded (through the context's pinned staging,
    # as the server's are) and then built; u64 keys are built in place.
    ctx = ph.Context(0)
    stream = hip_rt.Stream()
    warm = hip_rt.DeviceBuffer(32 << 20)  # the context's pinned chunks exist before timing, as a server's do
    ctx.copy_to_device(warm, np.zeros(32 << 20, np.uint8), stream)
    del warm
    def sequential_pass():
        bs, us = [], []
        for kind, n, payload, cfg in reqs:
            np_, B = ph.sizes(cfg, n)
            rc = ph.remap_capacity(cfg, n)
            dp, do, ds = (hip_rt.DeviceBuffer(max(np_ * B, 1) * 4), hip_rt.DeviceBuffer((np_ + 1) * 4),
                          hip_rt.DeviceBuffer(4))
            drm = hip_rt.DeviceBuffer(max(rc, 1) * 4) if rc else None
            dro = hip_rt.DeviceBuffer((np_ + 1) * 4) if rc else None
            outs = ph.device_outputs(dp, do, drm, dro)
            if kind == "bytes":
                dbytes = hip_rt.DeviceBuffer(max(payload[0].nbytes, 1))
                doffs = hip_rt.DeviceBuffer(payload[1].nbytes)
            stream.synchronize()
            t0 = time.perf_counter()
            if kind == "u64":
                ctx.build_u64_device(payload, n, cfg, outs, ds, stream)
                t1 = t0
            else:
                ctx.copy_to_device(dbytes, payload[0], stream)
                ctx.copy_to_device(doffs, payload[1], stream)
                t1 = time.perf_counter()
                ctx.build_bytes_device(dbytes, doffs, n, cfg, outs, ds, stream)
            stream.synchronize()
            bs.append(time.perf_counter() - t0)
            us.append(t1 - t0)
        return bs, us

    # two passes; each request's faster one is the sequential service time
    (b1, u1), (b2, u2) = sequential_pass(), sequential_pass()
    build_s = [min(x, y) for x, y in zip(b1, b2)]
    upload_s = [u if b == x else v for b, x, u, v in zip(build_s, b1, u1, u2)]
    done, lat = 0.0, []
    for a, b in zip(arrivals, build_s):
        done = max(done, a) + b
        lat.append(done - a)
    seq_span = done - arrivals[0]
    ctx.close()

    def serve(paced: bool):
        srv = ph.Server(slots=args.slots, hbm_budget=200 << 30)
        t0 = time.perf_counter()
        ids = []
        for (kind, n, payload, cfg), a in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if paced and a > now:
                time.sleep(a - now)
            ids.append(srv.submit_u64(payload, cfg, n) if kind == "u64" else
                       srv.submit_bytes(payload[0], payload[1], cfg))
        statuses = [srv.wait(i)[0] for i in ids]
        st = srv.stats()
        for i in ids:
            srv.release(i)
        srv.close()
        return st, statuses

    serve(False)  # warm-up pass, untimed: the device pool grows to the trace's working set
    # three passes each, paced at the arrivals and all at once (the server's
    # capacity); the pass of median span is reported
    def median_pass(paced: bool):
        runs = [serve(paced) for _ in range(3)]
        runs.sort(key=lambda r: r[0]["span_s"])
        return runs[1][0], [x for r in runs for x in r[1]], [total / r[0]["span_s"] / 1e6 for r in runs]

    st, statuses, paced_rates = median_pass(True)
    burst, b_statuses, burst_rates = median_pass(False)
    lat_ms = np.array(lat) * 1e3
    arrival_span = float(arrivals[-1] - arrivals[0])
    return {
        "workload": f"{R} requests (numpy seed 5): Poisson arrivals, mean 2000 req/s, 70% u64 log-U[1e4,1e8] keys "
                    "in HBM / 30% printable strings log-U[1e4,2e6] of length 8..64 in host memory, lambda {6,7,8}, "
                    f"alpha {{0.99,1}}, 200 GB HBM budget, {args.slots} slots",
        "requests": R, "keys": total,
        "p50_ms": round(st["p50_ms"], 2),
        "p99_ms": round(st["p99_ms"], 2),
        "served_mkeys_s": round(total / st["span_s"] / 1e6, 2),
        "served_mkeys_s_passes": [round(x, 1) for x in sorted(paced_rates)],
        "sequential_mkeys_s": round(total / seq_span / 1e6, 2),
        "speedup_vs_sequential": round(seq_span / st["span_s"], 2),
        "sequential_p50_ms": round(float(np.percentile(lat_ms, 50)), 2),
        "sequential_p99_ms": round(float(np.percentile(lat_ms, 99)), 2),
        "offered_mkeys_s": round(total / arrival_span / 1e6, 2),
        "sequential_upload_ms": round(sum(upload_s) * 1e3, 2),
        "sequential_build_ms": round((sum(build_s) - sum(upload_s)) * 1e3, 2),
        "grouped_launches": st["grouped_launches"],
        "grouped_requests": st["grouped_requests"],
        "burst": {"served_mkeys_s": round(total / burst["span_s"] / 1e6, 2),
                  "served_mkeys_s_passes": [round(x, 1) for x in sorted(burst_rates)],
                  "speedup_vs_sequential_builds": round(sum(build_s) / burst["span_s"], 2),
                  "p50_ms": round(burst["p50_ms"], 2), "p99_ms": round(burst["p99_ms"], 2),
                  "grouped_launches": burst["grouped_launches"], "grouped_requests": burst["grouped_requests"]},
        "failed": int(sum(1 for s in statuses + b_statuses if s)),
    }


def run_strings(args) -> dict:
    """Plain build of n printable random strings, lengths uniform in
    [8, 64] (numpy seed 6), bytes generated in HBM, lambda 6, P 2048,
    alpha 1."""
    n, lam, P, alpha = args.strings, 6.0, 2048, 1.0
    rng = np.random.default_rng(6)
    offs = np.zeros(n + 1, np.uint64)
    np.cumsum(rng.integers(8, 65, n, dtype=np.uint8), dtype=np.uint64, out=offs[1:])
    total_bytes = int(offs[-1])
    cfg = ph.make_config(lam, P, 0, alpha)
    np_, B = ph.sizes(cfg, n)
    ctx = ph.Context(0)
    apply_tuning_flags(ctx, args)
    stream = hip_rt.Stream()
    d_offs = hip_rt.DeviceBuffer.from_numpy(offs)
    d_bytes = hip_rt.DeviceBuffer(total_bytes)
    ph.gen_strings_device(6, n, d_offs, d_bytes, stream)
    dp, do, ds = hip_rt.DeviceBuffer(np_ * B * 4), hip_rt.DeviceBuffer((np_ + 1) * 4), hip_rt.DeviceBuffer(4)
    outs = ph.device_outputs(dp, do)
    stream.synchronize()
    for _ in range(max(1, args.warmup)):
        ctx.build_bytes_device(d_bytes, d_offs, n, cfg, outs, ds, stream)
    stream.synchronize()
    status = int(ds.to_numpy(np.uint32)[0])
    if status:
        raise RuntimeError(f"string build failed: {ph.STATUS_NAMES.get(ph.status_from_bits(status))}")
    steps = max(1, min(args.steps, 5))
    stage_sum: dict[str, float] = {}
    e0, e1 = hip_rt.Event(), hip_rt.Event()
    e0.record(stream)
    for _ in range(steps):
        ctx.build_bytes_device(d_bytes, d_offs, n, cfg, outs, ds, stream)
        for k, v in ctx.last_stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    e1.record(stream)
    stream.synchronize()
    status = int(ds.to_numpy(np.uint32)[0])
    if status:
        raise RuntimeError(f"string build failed: {ph.STATUS_NAMES.get(ph.status_from_bits(status))}")
    ms = e0.elapsed_ms(e1) / steps
    stages = {k: v / steps for k, v in stage_sum.items()}
    kernels = {k: v for k, v in stages.items() if k != "total"}
    key_bytes = total_bytes / n + 8
    res = {
        "workload": f"plain build of {n} printable random strings, lengths U[8,64] (numpy seed 6), bytes generated "
                    f"in HBM ({total_bytes / 1e9:.2f} GB), lambda {lam}, P {P}, alpha {alpha}",
        "n": n, "value": round(n / (ms / 1e3) / 1e6, 2), "unit": "Mkeys/s", "ms_per_step": round(ms, 3),
        "steps": steps, "mean_length": round(total_bytes / n, 2),
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "stages": stage_report(kernels, n, lam, P, alpha, key_bytes),
    }
    if not args.no_cpu_baseline:
        m = min(n, args.cpu_sample // 4)
        sample = d_bytes.to_numpy(np.uint8, int(offs[m]))
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po

        t0 = time.perf_counter()
        st = po.build_keys(po.hash_bytes(sample, offs[:m + 1], 0), lam, P)[0]
        dt = time.perf_counter() - t0
        if st != 0:
            raise RuntimeError(f"oracle baseline failed with status {st}")
        res["cpu_baseline"] = {
            "value": round(m / dt / 1e6, 3), "unit": "Mkeys/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
            "sample": f"the first {m} of the same strings: oracle XXH3 + build, OpenMP over partitions, {dt:.1f} s"}
    del d_offs, d_bytes, dp, do, ds
    ctx.close()
    return res


def run_host(args) -> dict:
    """The host-buffer build (phobic_build_u64): keys in pageable host memory,
    pilots, offsets and the remap table copied back to host memory, one
    synchronous call per build, as MPHFbuilder::build."""
    n, lam, P, alpha = args.n, args.lam, args.P, args.alpha
    cfg = ph.make_config(lam, P, 0, alpha)
    d = hip_rt.DeviceBuffer(n * 8)
    ph.gen_splitmix64_device(args.key_seed, n, d)
    hip_rt.synchronize()
    keys = d.to_numpy(np.uint64)
    del d
    ctx = ph.Context(0)
    r = ctx.build_u64(keys, cfg)  # warm-up: scratch, staging chunks
    times = []
    for _ in range(max(1, min(args.steps, 3))):
        t0 = time.perf_counter()
        r = ctx.build_u64(keys, cfg)
        times.append(time.perf_counter() - t0)
    best = min(times)
    ctx.close()
    return {"workload": f"phobic_build_u64 of {n} u64 keys from pageable host memory, outputs to host memory, "
                        f"lambda {lam}, P {P}, alpha {alpha}",
            "value": round(n / best / 1e6, 1), "unit": "Mkeys/s", "ms_per_build": round(best * 1e3, 2),
            "stages_ms": {k: round(v, 3) for k, v in r.times.items()}}


def bench_copy(args) -> int:
    """Copies of 1 GiB device to device, by the runtime (hipMemcpyAsync) and
    by the library's hand-written streaming kernel (phobic_copy_device): the
    known-size references for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE
    readings to HBM bytes."""
    nbytes = 1 << 30
    src, dst = hip_rt.DeviceBuffer(nbytes), hip_rt.DeviceBuffer(nbytes)
    src.zero()
    hip_rt.synchronize()
    rt = hip_rt.rt()
    e0, e1 = hip_rt.Event(), hip_rt.Event()
    res = {"mode": "copy", "bytes_read": nbytes, "bytes_written": nbytes}
    for name, fn in (("runtime", lambda: hip_rt.check(rt.hipMemcpyAsync(dst.ptr, src.ptr, nbytes, hip_rt.D2D, None),
                                                      "hipMemcpyAsync")),
                     ("kernel", lambda: ph.copy_device(dst, src, nbytes))):
        ms = []
        for _ in range(max(1, args.steps)):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_ms(e1))
        best = min(ms)
        res[name] = {"best_ms": round(best, 4), "gbs_read_plus_write": round(2 * nbytes / best / 1e6, 1)}
    print(json.dumps(res), flush=True)
    return 0


def part_line(metric: str, unit: str, res: dict, args, higher=True) -> dict:
    """One bench line for a part run alone (--mode batch / serve / strings)."""
    line = {"metric": metric, "value": res.get("value"), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8" if "strings" in metric else "u64", "data": "synthetic",
            "config": {"workload": res.get("workload")}}
    line.update({k: v for k, v in res.items() if k not in line})
    return line


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000_000)
    ap.add_argument("--lam", type=float, default=7.5)
    ap.add_argument("--P", type=int, default=2048)
    ap.add_argument("--alpha", type=float, default=0.99)
    ap.add_argument("--key-seed", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=64_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--mode", choices=["all", "plain", "strings", "batch", "serve", "copy", "host"], default="all")
    ap.add_argument("--strings", type=int, default=200_000_000, help="string count of the string-keyed build")
    ap.add_argument("--sets", type=int, default=4096)
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--slots", type=int, default=8, help="concurrent builds of the server")
    ap.add_argument("--autotune", type=int, default=0,
                    help="sample keys for the launch-shape autotuner (0 = default shapes)")
    ap.add_argument("--search-mode", type=int, default=None, help="pilot-search variant mask (0..3)")
    ap.add_argument("--search-waves", type=int, default=None, help="partitions per search workgroup")
    ap.add_argument("--route-per", type=int, default=None, help="records per thread of a level-1 routing tile")
    ap.add_argument("--max-fanout", type=int, default=None, help="bins per partitioning level (16..1024)")
    ap.add_argument("--split-per", type=int, default=None, help="records per thread of a split tile (0 = route-per)")
    ap.add_argument("--first-fanout", type=int, default=None, help="at most this many level-1 bins (0 = max-fanout)")
    ap.add_argument("--route-threads", type=int, default=None, help="threads of a routing workgroup (512 or 1024)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build path runs on a single device; --gpus {args.gpus} is not "
              "supported (sharding by partition range is described in DESIGN.md)", file=sys.stderr)
        return 2

    if args.mode == "copy":
        return bench_copy(args)
    if args.mode == "batch":
        res = run_batch(args)
        print(json.dumps(part_line("batched-vs-plain keys/s", "Mkeys/s", res, args)), flush=True)
        return 0 if res["identical_outputs"] and res["status"] == 0 else 1
    if args.mode == "serve":
        res = run_serve(args)
        res["value"] = res["p99_ms"]
        print(json.dumps(part_line("p99 latency", "ms", res, args, higher=False)), flush=True)
        return 0 if res["failed"] == 0 else 1
    if args.mode == "host":
        res = run_host(args)
        print(json.dumps(part_line("MPHF build Mkeys/s, host buffers", "Mkeys/s", res, args)), flush=True)
        return 0
    if args.mode == "strings":
        res = run_strings(args)
        print(json.dumps(part_line("MPHF build Mkeys/s, 200M strings", "Mkeys/s", res, args)), flush=True)
        return 0
    out = run_plain(args)
    if args.mode == "all":
        # the other figures of the metric, each on its own workload after the
        # headline's buffers are freed
        out["batched_vs_plain"] = run_batch(args)
        out["serving"] = run_serve(args)
        out["strings_200m"] = run_strings(args)
    print(json.dumps(out), flush=True)
    ok = out.get("bijection_violations", 0) == 0
    if args.mode == "all":
        ok = ok and out["batched_vs_plain"]["identical_outputs"] and out["serving"]["failed"] == 0
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
