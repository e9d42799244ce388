#!/usr/bin/env python3
"""Benchmark of the HIBF query path on one MI355X.

Workload (flat IBF): 8,192 technical bins x 2^23 rows (8 GiB bit matrix),
3 hash functions, built by the GPU bulk emplace of 2^17 generated uint64
values per bin; 1,000,000 queries x 250 values (70 % drawn from one bin's
inserted set, 30 % uniform); uint16 bulk_count.  A step is one batched
bulk_count launch over all queries, inputs resident in HBM.

Prints one JSON line: lookups/s (a lookup = one query value checked against
the IBF), the HBM roofline of the count kernel measured with HIP events on the
launch stream, the plain one-launch-per-query form, and a CPU baseline (the
oracle restatement on a bounded sample, on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--length", type=int, default=250)
    ap.add_argument("--bins", type=int, default=8192)
    ap.add_argument("--log2-bin-size", type=int, default=23)
    ap.add_argument("--hash-funs", type=int, default=3)
    ap.add_argument("--per-bin", type=int, default=1 << 17)
    ap.add_argument("--frac-in", type=float, default=0.7)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--plain-queries", type=int, default=2000)
    ap.add_argument("--check-queries", type=int, default=256)
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch-shape knob passed to hibf_gpu_tune_set (repeatable)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB elementwise copy kernel (known bytes) for counter calibration")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the query path runs on a "
              "single MI355X; sharding across GPUs is not part of this build).", file=sys.stderr)
        sys.exit(2)

    import torch
    from hibf_amd.device import DeviceIbf, as_device_u64
    from hibf_amd.synth import flat_ibf_queries

    from hibf_amd import native
    for kv in args.tune:
        k, v = kv.split("=")
        native.tune_set(k, int(v))

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    bins, bin_size, h = args.bins, 1 << args.log2_bin_size, args.hash_funs
    nq, L = args.queries, args.length

    log(f"build: {bins} TBs x 2^{args.log2_bin_size} rows, h={h}, {args.per_bin} values/bin")
    ibf = DeviceIbf(bins, bin_size, h)
    emplace_seed = args.seed
    t0 = time.perf_counter()
    ibf.emplace_synthetic(emplace_seed, 0, bins, args.per_bin)
    build_s = time.perf_counter() - t0
    log(f"emplace: {bins * args.per_bin} values in {build_s:.3f} s")

    log(f"queries: {nq} x {L}")
    values_h, offsets_h, _ = flat_ibf_queries(args.seed + 1, nq, L, bins, args.per_bin, args.frac_in, emplace_seed)
    values = as_device_u64(values_h)
    offsets = as_device_u64(offsets_h)
    counts = torch.empty((nq, bins), dtype=torch.uint16, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        ibf.bulk_count_into(values, offsets, counts, 2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: K steps; HIP events on the launch stream bracket each launch.
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in evs]
    ms_per_step = wall * 1e3 / args.steps
    lookups = nq * L
    value = lookups / (ms_per_step / 1e3)
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    row_bytes = h * ((bins + 63) // 64) * 8
    achieved = row_bytes * lookups / avg_kernel_s / 1e9
    log(f"batched: {ms_per_step:.2f} ms/step, kernel {np.mean(kernel_ms):.2f} ms, {value / 1e9:.3f} G lookups/s, "
        f"{achieved:.0f} GB/s of rows")

    # Plain form: one launch per query (async back to back), then per-query latency (synchronised).
    npl = min(args.plain_queries, nq)
    outs = counts[:1]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in range(npl):
        ibf.bulk_count_into(values, offsets[q:q + 2], outs, 2)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_value = npl * L / plain_s if npl else float("nan")
    lat = [0.0]
    for q in range(min(300, npl)):
        t1 = time.perf_counter()
        ibf.bulk_count_into(values, offsets[q:q + 2], outs, 2)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    log(f"plain: {plain_value / 1e9:.4f} G lookups/s; latency p50 {np.percentile(lat, 50):.3f} ms "
        f"p99 {np.percentile(lat, 99):.3f} ms")

    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.int64, device=dev)
        b = torch.empty_like(a)
        torch.add(a, 0, out=b)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del a, b

    # Parity on a bounded sample (oracle on the downloaded matrix) and the CPU baseline.
    cpu = None
    parity = None
    if not args.skip_cpu or args.check_queries > 0:
        import oracle as O
        words = ibf.download()
        ref = O.Ibf.from_words(bins, bin_size, h, words)
        threads = max(1, min(16, os.cpu_count() or 1))
        rng = np.random.default_rng(args.seed + 7)
        sample = np.sort(rng.choice(nq, size=min(args.check_queries, nq), replace=False))
        step()
        torch.cuda.synchronize()
        idx = torch.as_tensor(sample.astype(np.int64), device=dev)
        got = counts.view(torch.int16).index_select(0, idx).cpu().numpy().view(np.uint16)
        sv = np.concatenate([values_h[offsets_h[q]:offsets_h[q + 1]] for q in sample])
        so = np.arange(len(sample) + 1, dtype=np.uint64) * np.uint64(L)
        want = ref.bulk_count_batch(sv, so, 2, threads)
        parity = bool(np.array_equal(got, want))
        log(f"parity vs oracle on {len(sample)} queries: {parity}")
        if not args.skip_cpu:
            probe = 64
            t1 = time.perf_counter()
            ref.bulk_count_batch(values_h[: probe * L], offsets_h[: probe + 1], 2, threads)
            rate = probe / max(time.perf_counter() - t1, 1e-6)
            n_cpu = int(min(nq, max(probe, rate * args.cpu_seconds)))
            t1 = time.perf_counter()
            ref.bulk_count_batch(values_h[: n_cpu * L], offsets_h[: n_cpu + 1], 2, threads)
            cpu_s = time.perf_counter() - t1
            cpu = {"value": n_cpu * L / cpu_s, "unit": "lookups/s", "cores": threads, "kind": "port",
                   "sample": f"first {n_cpu} queries x {L} values of the same workload, oracle bulk_count "
                             f"(uint16) with {threads} OpenMP threads, one agent per thread, {cpu_s:.1f} s"}
            log(f"cpu baseline: {cpu['value'] / 1e6:.2f} M lookups/s on {threads} threads")
        del words, ref

    out = {
        "metric": "k-mer lookups/s (batched IBF bulk_count)",
        "value": value,
        "unit": "lookups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 hash / u16 counters",
        "data": "synthetic (seeded generator values, uniform uint64 noise)",
        "config": {
            "workload": f"flat IBF: {bins} technical bins x 2^{args.log2_bin_size} bits, h={h}, "
                        f"{args.per_bin} values/bin by GPU bulk emplace; {nq} queries x {L} values "
                        f"({int(args.frac_in * 100)}% from one bin); uint16 bulk_count",
            "bins": bins, "bin_size": bin_size, "hash_functions": h, "queries": nq, "query_length": L,
            "counter": "uint16",
            "tuning": {k: native.tune_get(k) for k in ("count.wide_kernel", "count.wide_unroll",
                                                        "count.wide_waves")},
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel_ms": float(np.mean(kernel_ms)), "bytes_per_lookup": row_bytes,
        },
        "cpu_baseline": cpu,
        "plain_per_query": {"value": plain_value, "unit": "lookups/s", "queries": npl,
                            "latency_ms_p50": float(np.percentile(lat, 50)),
                            "latency_ms_p99": float(np.percentile(lat, 99))},
        "batched_vs_plain": value / plain_value,
        "parity_sample_ok": parity,
        "build_emplace_s": build_s,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
