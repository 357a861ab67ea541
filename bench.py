"""Benchmark: plain PHOBIC build throughput on one MI355X.

One step = one full device build (hash -> partition -> bucket sort ->
pilot search) of n distinct synthetic u64 keys that are already resident in
HBM, producing pilots u32[np*B] and partition offsets u32[np+1] in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]

Prints one JSON line (metric, value in Mkeys/s, per-stage device times, the
dominant kernel's HBM roofline figure and a CPU baseline of the oracle on a
bounded sample of the same workload).
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


def stage_bytes_per_key(name: str, lam: float, P: int) -> float:
    """Algorithmic bytes per key of each kernel: the rows of the reference's
    stage structure that kernel performs (per-key figures: S1 8+1+4/l,
    S2 12/l+1028/P, S3 8/l+8/P, S4 8+1+4/l+8, S5 8+1028/P+12/l+4/l)."""
    s1 = 8 + 1 + 4 / lam
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


def cpu_baseline(n_sample: int, lam: float, P: int, kseed: int) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    keys = po.splitmix64(kseed, n_sample)
    t0 = time.perf_counter()
    hk = po.hash_u64(keys, 0)
    st, _, _ = po.build_keys(hk, lam, P)
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
                  f"build (bucketing, bucket sort, pilot search), OpenMP over partitions, {dt:.1f} s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000_000)
    ap.add_argument("--lam", type=float, default=7.5)
    ap.add_argument("--P", type=int, default=2048)
    ap.add_argument("--alpha", type=float, default=1.0)
    ap.add_argument("--key-seed", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=16_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build path runs on a single device; --gpus {args.gpus} is not "
              "supported (sharding by partition range is described in DESIGN.md)", file=sys.stderr)
        return 2

    n, lam, P = args.n, args.lam, args.P
    alpha = args.alpha
    cfg = ph.make_config(lam, P, 0, alpha)
    np_, B = ph.sizes(cfg, n)
    rcap = ph.remap_capacity(cfg, n)
    ctx = ph.Context(0)
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
    dom = max(kernels, key=kernels.get)
    bpk = stage_bytes_per_key(dom, lam, P)
    achieved = bpk * n / (kernels[dom] / 1e3) / 1e9
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
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 3),
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_key": round(bpk, 3),
        },
    }
    if not args.no_verify:
        out_pos = hip_rt.DeviceBuffer(n * 4)
        ctx.eval_u64_device(dk, n, cfg, np_, mphf, out_pos, stream)
        bm, bad = hip_rt.DeviceBuffer((n + 31) // 32 * 4), hip_rt.DeviceBuffer(8)
        ctx.check_bijection_device(out_pos, n, n, bm, bad, stream)
        stream.synchronize()
        out["bijection_violations"] = int(bad.to_numpy(np.uint64)[0])
        del out_pos, bm, bad
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, lam, P, args.key_seed)
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
