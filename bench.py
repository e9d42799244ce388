#!/usr/bin/env python3
"""Benchmark of the HIBF query path on one MI355X.

Default workload (the batched form on the 32,768-user-bin HIBF): user bins
of log-normal sizes (median 1e5, sigma 1.5, clipped to [1e3, 5e7]), tmax 256,
2 hash functions, FPR 0.05 / relaxed 0.3, on the reference's default layout
(HyperLogLog estimates, union estimates, rearrangement), built on the device;
2,000,000 queries of log-normal length (median 300, clipped to [50, 5000]),
60 % of their values from one user bin, threshold uint16_t(0.7 L), answered by
membership_for in hipGraph-captured batches of 65,536 queries.  A step is one
pass over all 2M queries, inputs resident in HBM.

Prints one JSON line: lookups/s (a lookup = one query value checked against
one visited IBF), the HBM roofline of the descent kernel (HIP events on the
launch stream), the plain one-launch-per-query form, the oracle on this
host's cores as the CPU baseline, the serving loop's p50/p99 at 50k and 500k
queries/s, and the flat 8,192-bin IBF bulk_count as an extra leg.
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# Random 32-byte row reads from a 16 GiB buffer, one 128-byte line each: the fastest of the load forms and
# launch shapes measured, 5.36e10 rows/s with non-temporal loads, two lanes per row
# (profiles/random_row_gather_ceiling.md; cached loads reach 4.89e10).
RANDOM_LINE_CEILING = 5.36e10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--serve-rates", default="50000,100000,200000,500000",
                    help="Poisson arrival rates (queries/s) of the serving workload")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-host-gb", type=float, default=200.0,
                    help="serve: skip offered rates whose generated queries would need more host memory than this")
    ap.add_argument("--serve-windows", default="200,1000,2000", help="batching windows (us)")
    ap.add_argument("--serve-pageable", action="store_true",
                    help="hand the serving loop pageable query values (staged) instead of a pinned receive buffer")
    ap.add_argument("--genomes", type=int, default=4096,
                    help="serve_reads: random genomes (user bins) of --genome-length bases")
    ap.add_argument("--genome-length", type=int, default=20000)
    ap.add_argument("--serve-minimisers", action="store_true",
                    help="serve_reads: hash genomes and reads to minimisers (ungapped 19, window 23, seqan3's seed) "
                         "instead of 19-mers")
    ap.add_argument("--workload", choices=["hibf", "flat", "serve", "serve_reads"], default="hibf",
                    help="hibf (default): 32,768-user-bin HIBF membership (log-normal batched hipGraph form as the "
                         "value, fixed-length batch, plain per-query form, serving p50/p99, flat IBF leg); "
                         "flat: 8,192-bin flat IBF bulk_count alone")
    ap.add_argument("--no-flat-leg", action="store_true", help="hibf: skip the flat IBF leg")
    ap.add_argument("--no-serve-leg", action="store_true", help="hibf: skip the serving leg")
    ap.add_argument("--leg-serve-rates", default="50000,500000",
                    help="hibf: offered rates (queries/s) of the serving leg, window --leg-serve-window")
    ap.add_argument("--leg-serve-window", type=float, default=200.0, help="hibf: serving leg batching window (us)")
    ap.add_argument("--leg-serve-seconds", type=float, default=1.0)
    ap.add_argument("--hibf-queries", type=int, default=500_000)
    ap.add_argument("--hibf-batched-queries", type=int, default=2_000_000)
    ap.add_argument("--reads", type=int, default=200_000,
                    help="250-base reads for the k-mer hashing leg of the HIBF workload (0 skips it)")
    ap.add_argument("--hibf-count-queries", type=int, default=50_000,
                    help="queries of (a) for the HIBF bulk_count leg (uint16 counts of every user bin, t = 0.7 L; "
                         "0 skips it)")
    ap.add_argument("--graph-batch", type=int, default=65536)
    ap.add_argument("--streams", type=int, default=3,
                    help="hibf: streams the graph batches of (b) alternate over (one agent each), so that a batch "
                         "starts while the previous ones drain")
    ap.add_argument("--hibf-user-bins", type=int, default=32768)
    ap.add_argument("--recompute-layout", action="store_true",
                    help="hibf: compute the layout (union estimates + rearrangement on the device) instead of "
                         "reading the stored one, and report its times")
    ap.add_argument("--layout-only", action="store_true",
                    help="hibf: stop after the sketches and the layout")
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
    ap.add_argument("--contains-values", type=int, default=1 << 20,
                    help="values of the bulk_contains leg of the flat workload (0 skips it)")
    ap.add_argument("--membership-queries", type=int, default=1_000_000,
                    help="queries of the membership_for leg of the flat workload, t = 0.7 L (0 skips it)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="launch-shape knob passed to hibf_gpu_tune_set (repeatable)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes instead of autotuning them on a sub-batch")
    ap.add_argument("--hibf-tune-queries", type=int, default=0,
                    help="hibf: queries of (b) the launch shapes are tuned on (default: two graph batches per stream)")
    ap.add_argument("--autotune-queries", type=int, default=100_000,
                    help="flat: queries of the workload the autotuner times each candidate shape on")
    ap.add_argument("--autotune-save", default=None, metavar="PATH", help="write the autotuner record as JSON")
    ap.add_argument("--traffic-sweep", action="store_true",
                    help="hibf: after the timed region, launch (a) and every batch of (b) once under each launch shape "
                         "of the tuner's space (for rocprofv3 counter runs; a 1-element fill kernel marks each shape's (a) and (b))")
    ap.add_argument("--sweep-rounds", type=int, default=1, help="hibf: rounds of --traffic-sweep (best time kept)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB elementwise copy kernel (known bytes) for counter calibration")
    return ap.parse_args()


def recorded_traffic(key: str, args, layout_text: str | None = None) -> float | None:
    """HBM bytes per launch measured with rocprofv3 counters on this exact
    workload (profiles/traffic.json), or None when the workload (or, for an
    HIBF, the layout or the launch shape the record was measured with)
    differs."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)[key]
    except (OSError, KeyError, ValueError):
        return None
    if any(getattr(args, k) != v for k, v in rec["workload"].items()):
        return None
    if "layout_sha1" in rec and (layout_text is None or
                                 hashlib.sha1(layout_text.encode()).hexdigest()[:16] != rec["layout_sha1"]):
        return None
    if "shapes" in rec:  # one row per launch shape (a shape can change the line count): the current one's
        from hibf_amd import native
        nk = len(rec["shape_keys"])
        now = [native.tune_get(k) for k in rec["shape_keys"]]
        row = next((r for r in rec["shapes"] if r[:nk] == now), None)
        if row is None:
            return None
        rec = {**rec, **dict(zip(rec["shape_columns"], row[nk:]))}
    return (rec["fetch_kib"] * rec["fetch_factor"] + rec["write_kib"] * rec["write_factor"]) * 1024.0


def host_cpu(simd: bool) -> dict:
    """CPU model and core count of the host the CPU baseline runs on, and the
    counting path the oracle used."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    path = ("AVX-512BW (counting_vector.hpp:252-278 restated: masked add of 32 uint16 counters per register, the "
            "path of a -march=native reference build on this host)" if simd else
            "scalar (countr_zero over the AND of the rows, counting_vector.hpp:279-311; this host has no AVX-512BW)")
    return {"cpu_model": model, "nproc": os.cpu_count(), "counting_path": path}


def timed_sample(run, n_probe: int, n_max: int, seconds: float) -> tuple[int, float]:
    """Runs run(n) on a probe of n_probe units, then on as many units as fit
    `seconds`; returns (units, wall seconds) of the second run."""
    t1 = time.perf_counter()
    run(n_probe)
    rate = n_probe / max(time.perf_counter() - t1, 1e-6)
    n = int(min(n_max, max(n_probe, rate * seconds)))
    t1 = time.perf_counter()
    run(n)
    return n, time.perf_counter() - t1


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (the query path runs on a "
              "single MI355X; sharding across GPUs is not part of this build).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "flat":
        run_flat(args)
    elif args.workload == "hibf":
        run_hibf(args)
    elif args.workload == "serve_reads":
        run_serve_reads(args)
    else:
        run_serve(args)


def run_flat(args):
    print(json.dumps(flat_measure(args)), flush=True)


def flat_measure(args, leg: bool = False):
    """The flat IBF workload; returns its JSON record.  leg=True: as an extra
    leg of the HIBF run (library default shapes, shorter CPU sample)."""
    import torch
    from hibf_amd.device import Agent, DeviceIbf, as_device_u64, to_numpy
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

    tuned = None
    if not args.no_autotune and not args.tune and not leg:
        from hibf_amd import autotune
        nt = min(args.autotune_queries, nq)
        rec = autotune.tune_flat_count(ibf, values[: int(offsets_h[nt])], offsets[: nt + 1], 2, reps=2)
        tuned = {"best": rec["best"], "best_ms": rec["best_ms"], "default_ms": rec["default_ms"],
                 "queries": nt, "candidates": len(rec["candidates"])}
        log(f"autotune ({len(rec['candidates'])} shapes on {nt} queries): {rec['best']} "
            f"{rec['best_ms']:.3f} ms (default {rec['default_ms']:.3f} ms)")
        if args.autotune_save:
            rec["command"] = "python bench.py " + " ".join(sys.argv[1:])
            autotune.save_record(rec, args.autotune_save)

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

    # Plain form: one launch per query (async back to back), then per-query latency (synchronised).  The
    # launches call the C-ABI directly, as a C++ caller would, without the Python wrapper's buffer checks.
    npl = min(args.plain_queries, nq)
    outs = counts[:1]
    c_count = native.lib().hibf_gpu_ibf_bulk_count
    p_vals, p_out, p_s = ctypes.c_void_p(values.data_ptr()), ctypes.c_void_p(outs.data_ptr()), \
        ctypes.c_void_p(stream.cuda_stream)
    p_off = [ctypes.c_void_p(offsets[q:q + 2].data_ptr()) for q in range(npl)]

    def plain_count(q):
        native.check(c_count(ibf.h, p_vals, p_off[q], 1, 2, p_out, p_s))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in range(npl):
        plain_count(q)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_value = npl * L / plain_s if npl else float("nan")
    lat = [0.0]
    for q in range(min(300, npl)):
        t1 = time.perf_counter()
        plain_count(q)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    log(f"plain: {plain_value / 1e9:.4f} G lookups/s; latency p50 {np.percentile(lat, 50):.3f} ms "
        f"p99 {np.percentile(lat, 99):.3f} ms")

    # bulk_contains: one result row (bin_words words) per value; reads h rows, writes one.
    contains = None
    nc = min(args.contains_values, values.numel())
    if nc > 0:
        cvals = values[:nc]
        cout = torch.empty((nc, (bins + 63) // 64), dtype=torch.uint64, device=dev)
        ibf.bulk_contains_into(cvals, cout)
        torch.cuda.synchronize()
        cev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
        for a, b in cev:
            a.record(stream)
            ibf.bulk_contains_into(cvals, cout)
            b.record(stream)
        torch.cuda.synchronize()
        c_ms = float(np.mean([a.elapsed_time(b) for a, b in cev]))
        c_bytes = (h + 1) * ((bins + 63) // 64) * 8
        contains = {"value": nc / (c_ms / 1e3), "unit": "values/s", "values": nc, "kernel_ms": c_ms,
                    "bytes_per_value": c_bytes, "achieved_GBs": c_bytes * nc / (c_ms / 1e3) / 1e9}
        log(f"bulk_contains: {nc} values in {c_ms:.3f} ms, {contains['value'] / 1e6:.1f} M values/s, "
            f"{contains['achieved_GBs']:.0f} GB/s")

    # membership_for over the whole batch at t = 0.7 L: count, threshold pass, scan, compaction (per chunk of
    # queries whose uint16 counts fit the agent's 256 MiB buffer).  Wall clock: the call returns the hit total.
    member = None
    nm = min(args.membership_queries, nq)
    if nm > 0:
        m_thr = int(0.7 * L)
        m_off = offsets[: nm + 1]
        agent = Agent()
        ibf.membership_for(values, m_off, m_thr, agent)
        m_s = []
        for _ in range(2):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            m_ids, m_hoff = ibf.membership_for(values, m_off, m_thr, agent)
            torch.cuda.synchronize()
            m_s.append(time.perf_counter() - t1)
        m_best = min(m_s)
        member = {"value": nm * L / m_best, "unit": "lookups/s", "queries": nm, "threshold": m_thr,
                  "seconds": m_best, "hits": int(m_ids.numel()), "vs_bulk_count": (nm * L / m_best) / value}
        log(f"membership_for: {nm} queries in {m_best * 1e3:.1f} ms, {member['value'] / 1e9:.3f} G lookups/s, "
            f"{member['hits']} hits")

    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.int64, device=dev)
        b = torch.empty_like(a)
        torch.add(a, 0, out=b)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del a, b

    # Parity on a bounded sample (oracle on the downloaded matrix) and the CPU baseline.
    cpu = None
    parity = None
    emplace_ok = None
    if not args.skip_cpu or args.check_queries > 0:
        import oracle as O
        from hibf_amd.synth import synth_values
        words = ibf.download()
        # The GPU emplace against the oracle's own emplace of the same generated values, on 64 sampled bins
        # (bins are independent columns; bin 0 and the last bin included): their bit columns must be equal
        # before the downloaded matrix serves as the oracle's filter below.
        rng = np.random.default_rng(args.seed + 7)
        col_bins = np.unique(np.concatenate([[0, bins - 1], rng.choice(bins, size=62, replace=False)]))
        own = O.Ibf(bins, bin_size, h)
        for b in col_bins:
            own.emplace(synth_values(emplace_seed, int(b), np.arange(args.per_bin, dtype=np.uint64)), int(b))
        bw = (bins + 63) // 64
        wd, od = words.reshape(bin_size, bw), own.words.reshape(bin_size, bw)
        emplace_ok = all(np.array_equal((wd[:, b // 64] >> np.uint64(b % 64)) & np.uint64(1),
                                        (od[:, b // 64] >> np.uint64(b % 64)) & np.uint64(1)) for b in col_bins)
        del own, od
        log(f"emplace columns of {col_bins.size} bins equal to the oracle's emplace: {emplace_ok}")
        ref = O.Ibf.from_words(bins, bin_size, h, words)
        threads = max(1, min(16, os.cpu_count() or 1))
        sample = np.sort(rng.choice(nq, size=min(args.check_queries, nq), replace=False))
        step()
        torch.cuda.synchronize()
        idx = torch.as_tensor(sample.astype(np.int64), device=dev)
        got = counts.view(torch.int16).index_select(0, idx).cpu().numpy().view(np.uint16)
        sv = np.concatenate([values_h[offsets_h[q]:offsets_h[q + 1]] for q in sample])
        so = np.arange(len(sample) + 1, dtype=np.uint64) * np.uint64(L)
        want = ref.bulk_count_batch(sv, so, 2, threads)
        parity = bool(np.array_equal(got, want)) and emplace_ok
        if contains is not None:
            pick = rng.choice(nc, size=min(256, nc), replace=False)
            got_c = cout.index_select(0, torch.as_tensor(pick.astype(np.int64), device=dev)).cpu().numpy()
            want_c = np.stack([ref.bulk_contains(int(values_h[i])) for i in pick])
            contains["parity_sample_ok"] = bool(np.array_equal(got_c, want_c))
            parity = parity and contains["parity_sample_ok"]
        if member is not None:
            hoff = to_numpy(m_hoff)
            ids = to_numpy(m_ids)
            ok = int(hoff[nm]) == ids.size
            in_m = sample[sample < nm]
            for j, q in enumerate(in_m):
                got_ids = ids[int(hoff[q]):int(hoff[q + 1])].tolist()
                want_ids = np.nonzero(want[int(np.searchsorted(sample, q))] >= m_thr)[0].tolist()
                if j < 16:  # the oracle's own membership_for on a few of them
                    ok = ok and ref.membership_for(values_h[offsets_h[q]:offsets_h[q + 1]], m_thr) == want_ids
                ok = ok and got_ids == want_ids
            member["parity_sample_ok"] = bool(ok)
            parity = parity and member["parity_sample_ok"]
        log(f"parity vs oracle on {len(sample)} queries: {parity}")
        if not args.skip_cpu:
            def run_cpu(n, t):
                ref.bulk_count_batch(values_h[: n * L], offsets_h[: n + 1], 2, t)
            cpu_budget = args.cpu_seconds / (4 if leg else 1)
            simd = O.set_simd_counting(True)
            n_cpu, cpu_s = timed_sample(lambda n: run_cpu(n, threads), 64, nq, cpu_budget)
            n_one, one_s = timed_sample(lambda n: run_cpu(n, 1), 16, nq, cpu_budget / 4)
            O.set_simd_counting(False)
            cpu = {"value": n_cpu * L / cpu_s, "unit": "lookups/s", "cores": threads, "kind": "port",
                   "sample": f"first {n_cpu} queries x {L} values of the same workload, oracle bulk_count "
                             f"(uint16) with {threads} OpenMP threads, one agent per thread, {cpu_s:.1f} s",
                   "one_thread": {"value": n_one * L / one_s, "queries": n_one, "seconds": one_s}, **host_cpu(simd)}
            log(f"cpu baseline: {cpu['value'] / 1e6:.2f} M lookups/s on {threads} threads, "
                f"{cpu['one_thread']['value'] / 1e6:.2f} M lookups/s on 1 thread")
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
                                                        "count.wide_waves", "count.counter_bits",
                                                        "count.wide_nt_loads")},
            "autotune": tuned,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": recorded_traffic("flat", args),
            "kernel_ms": float(np.mean(kernel_ms)), "bytes_per_lookup": row_bytes,
        },
        "cpu_baseline": cpu,
        "plain_per_query": {"value": plain_value, "unit": "lookups/s", "queries": npl,
                            "latency_ms_p50": float(np.percentile(lat, 50)),
                            "latency_ms_p99": float(np.percentile(lat, 99))},
        "batched_vs_plain": value / plain_value,
        "bulk_contains": contains,
        "membership_for": member,
        "parity_sample_ok": parity,
        "emplace_columns_ok": emplace_ok,
        "build_emplace_s": build_s,
    }
    del ibf, values, offsets, counts
    torch.cuda.empty_cache()
    return out


def hibf_sketch_leg(sizes, spec, args):
    """HyperLogLog sketches (12 bits, the config default) of the synthetic user
    bins on the device, as the reference's build computes them before the
    layout (compute_sketches + estimate_kmer_counts).  Returns the leg's record
    (None without args) and the truncated estimates."""
    import torch
    from hibf_amd.device import hll_estimates, hll_sketches_synthetic, to_numpy
    bits = 12
    hll_sketches_synthetic(sizes[:64], spec["build_seed"], bits)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    regs = hll_sketches_synthetic(sizes, spec["build_seed"], bits)
    sketch_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    est = hll_estimates(regs)
    est_s = time.perf_counter() - t0
    counts = est.astype(np.uint64)
    total = float(sizes.sum())
    log(f"(f) sketches: {sizes.size} user bins, {total:.3e} values in {sketch_s * 1e3:.1f} ms "
        f"({total / sketch_s / 1e9:.1f} G values/s), estimates {est_s * 1e3:.1f} ms; "
        f"median relative error {np.median(np.abs(counts / sizes - 1)):.4f}")
    if args is None:
        return None, counts, regs
    leg = {"user_bins": int(sizes.size), "values": total, "bits": bits, "seconds": sketch_s,
           "values_per_s": total / sketch_s, "estimate_seconds": est_s,
           "median_relative_error": float(np.median(np.abs(counts / sizes - 1)))}
    if args.check_queries > 0 or not args.skip_cpu:
        import oracle as O
        from hibf_amd.synth import synth_values
        rng = np.random.default_rng(args.seed + 31)
        small = np.nonzero(sizes <= 1_000_000)[0]
        pick = rng.choice(small, size=min(12, small.size), replace=False)
        host = to_numpy(regs[torch.as_tensor(pick.astype(np.int64), device=regs.device)])
        ok = True
        for row, u in zip(host, pick):
            want = O.hll_sketch(synth_values(spec["build_seed"], int(u), np.arange(int(sizes[u]), dtype=np.uint64)),
                                bits)
            ok = ok and np.array_equal(row, want) and float(est[u]) == O.hll_estimate(want)
        leg["parity_sample_ok"] = bool(ok)
        if not args.skip_cpu:
            u = int(pick[0])
            vals = synth_values(spec["build_seed"], u, np.arange(int(sizes[u]), dtype=np.uint64))
            reps = max(1, int(2e7 // max(vals.size, 1)))
            t0 = time.perf_counter()
            for _ in range(reps):
                O.hll_sketch(vals, bits)
            cpu_s = time.perf_counter() - t0
            leg["cpu_baseline"] = {"value": reps * vals.size / cpu_s, "unit": "values/s", "cores": 1, "kind": "port",
                                   "sample": f"oracle hyperloglog add of user bin {u} ({vals.size} values) x {reps}"}
    return leg, counts, regs


def run_hibf(args):
    """32,768-user-bin HIBF: (b) 2M queries of log-normal length in
    hipGraph-captured batches of 64k (the reported value); (a) 500k x 250
    queries, t = 0.7 L, one persistent launch; (c) the plain
    one-launch-per-query form on a sample of (b); (d) bulk_count; (e) k-mer
    hashing of reads; (f) the build-side sketches; then the serving loop at
    50k and 500k queries/s and the flat 8,192-bin IBF as extra legs."""
    import torch
    from hibf_amd import autotune, native
    from hibf_amd import workloads as W
    from hibf_amd.device import Agent, DeviceHibf, as_device_u64

    for kv in args.tune:
        k, v = kv.split("=")
        native.tune_set(k, int(v))
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    spec = W.HIBF_AT_SCALE
    n_ub = args.hibf_user_bins
    sizes = W.hibf_user_bin_sizes(spec["size_seed"], n_ub)
    # (f) build side: one HyperLogLog sketch per user bin on the device; the layout sees their estimates.
    sketch_leg, counts, regs = hibf_sketch_leg(sizes, spec, args)
    t0 = time.perf_counter()
    layout_times = {}
    layout = W.cached_layout(counts, spec["config"], registers=regs, times=layout_times,
                             reuse=not args.recompute_layout, write=not args.recompute_layout)
    layout_s = time.perf_counter() - t0
    del regs
    log(f"layout ({n_ub} user bins, {sizes.sum():.3e} values, HyperLogLog estimates, union estimates and "
        f"rearrangement): {layout_s:.1f} s" + (f" {layout_times}" if layout_times else " (stored)"))
    if args.layout_only:
        print(json.dumps({"metric": "HIBF layout (union estimates + rearrangement)", "value": layout_s, "unit": "s",
                          "n_gpus": 1, "higher_is_better": False, "config": {"workload": f"layout of {n_ub} user bins"},
                          "layout": layout_times or None, "sketches": sketch_leg}), flush=True)
        return
    t0 = time.perf_counter()
    hibf = DeviceHibf.build_synthetic(sizes, spec["build_seed"], layout_text=layout, **spec["config"])
    build_s = time.perf_counter() - t0
    tab = hibf.tables()
    log(f"build: {len(tab['geoms'])} IBFs, {tab['n_words'] * 8 / 1e9:.2f} GB, depth {tab['depth']}, {build_s:.1f} s")
    agent = Agent()
    stream = torch.cuda.current_stream()

    def thresholds_of(offsets_h):
        lens = np.diff(offsets_h.astype(np.int64))
        return np.minimum((0.7 * lens).astype(np.int64), 65535).astype(np.uint16)

    def upload_u16(a):
        return torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).to(dev).view(torch.uint16)

    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.int64, device=dev)
        b = torch.empty_like(a)
        torch.add(a, 0, out=b)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del a, b

    # (a) fixed-length batch, one persistent launch per step.
    nq = args.hibf_queries
    lengths = W.query_lengths(args.seed + 3, nq, fixed=250)
    va_h, oa_h, _ = W.hibf_queries(args.seed + 4, sizes, spec["build_seed"], lengths, 0.8)
    ta_h = thresholds_of(oa_h)
    va, oa, ta = as_device_u64(va_h), as_device_u64(oa_h), upload_u16(ta_h)

    # (b)'s queries: log-normal lengths, answered in hipGraph batches below.
    nb = args.hibf_batched_queries
    lengths_b = W.query_lengths(args.seed + 5, nb)
    vb_h, ob_h, _ = W.hibf_queries(args.seed + 6, sizes, spec["build_seed"], lengths_b, 0.6)
    tb_h = thresholds_of(ob_h)
    vb, ob, tb = as_device_u64(vb_h), as_device_u64(ob_h), upload_u16(tb_h)

    # (b)'s graph batches alternate over NS streams, batch b on stream b % NS with its own agent (workspace,
    # counters): consecutive batches are independent, so a batch's blocks start on the CUs the previous
    # batches' last queries leave free.
    G = args.graph_batch
    NS = max(1, args.streams)
    agents_b = [agent] + [Agent() for _ in range(NS - 1)]
    streams_b = [stream] + [torch.cuda.Stream() for _ in range(NS - 1)]

    # Launch shapes tuned on the first queries of (b), the workload of the reported value, in the form the
    # timed batches run in: batches of G launched over the NS streams, timed together.
    tuned = None
    if not args.no_autotune and not args.tune:
        from hibf_amd import autotune
        nt = min(args.hibf_tune_queries or 2 * NS * G, nb)
        ids_t = torch.empty(8 * nt, dtype=torch.uint64, device=dev)
        hs_t = torch.empty(nt, dtype=torch.uint64, device=dev)
        hc_t = torch.empty(nt, dtype=torch.uint64, device=dev)
        rec = autotune.tune_hibf_membership_batches(hibf, agents_b, streams_b, vb, ob[: nt + 1], tb[:nt], ids_t,
                                                    hs_t, hc_t, G, reps=3)
        del ids_t, hs_t, hc_t
        nbt = (nt + G - 1) // G
        tuned = {"best": rec["best"], "best_ms": rec["best_ms"], "default_ms": rec["default_ms"], "queries": nt,
                 "workload": f"the first {nbt} graph batches of (b) over {NS} streams, launch form",
                 "candidates": len(rec["candidates"])}
        log(f"autotune ({len(rec['candidates'])} shapes on the first {nbt} batches of (b) over {NS} streams): "
            f"{rec['best']} {rec['best_ms']:.3f} ms (default {rec['default_ms']:.3f} ms)")
        if args.autotune_save:
            rec["command"] = "python bench.py " + " ".join(sys.argv[1:])
            autotune.save_record(rec, args.autotune_save)

    cap = 8 * nq
    ids = torch.empty(cap, dtype=torch.uint64, device=dev)
    hs = torch.empty(nq, dtype=torch.uint64, device=dev)
    hc = torch.empty(nq, dtype=torch.uint64, device=dev)
    hibf.membership_prepare(agent, nq)
    hibf.membership_launch(agent, va, oa, ta, ids, hs, hc)
    torch.cuda.synchronize()
    total, dropped = agent.status()
    assert not dropped, "output capacity too small"
    agent.stats(reset=True)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(stream)
        hibf.membership_launch(agent, va, oa, ta, ids, hs, hc)
        b.record(stream)
    torch.cuda.synchronize()
    wall_a = (time.perf_counter() - t0) / args.steps
    kern_a = float(np.mean([a.elapsed_time(b) for a, b in evs])) / 1e3
    st = agent.stats(reset=True)
    lookups_a = st["lookups"] / args.steps
    bytes_a = st["row_bytes"] / args.steps
    log(f"(a) {nq} x 250: {wall_a * 1e3:.2f} ms/step, {lookups_a / wall_a / 1e9:.3f} G lookups/s, "
        f"{bytes_a / kern_a / 1e9:.0f} GB/s rows, {st['visits'] / args.steps / nq:.2f} IBFs/query, "
        f"{total / nq:.3f} hits/query")

    nbatches = (nb + G - 1) // G
    cap_b = 8 * G
    ids_b = torch.empty(nbatches * cap_b, dtype=torch.uint64, device=dev)
    hs_b = torch.empty(nb, dtype=torch.uint64, device=dev)
    hc_b = torch.empty(nb, dtype=torch.uint64, device=dev)
    for ag in agents_b:
        hibf.membership_prepare(ag, G)
    graphs = []
    for b in range(nbatches):
        q0, q1 = b * G, min(nb, (b + 1) * G)
        args_b = (agents_b[b % NS], vb, ob[q0:q1 + 1], tb[q0:q1], ids_b[b * cap_b:(b + 1) * cap_b], hs_b[q0:q1],
                  hc_b[q0:q1])
        hibf.membership_launch(*args_b)  # warm the capture inputs once outside the graph
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            hibf.membership_launch(*args_b)
        graphs.append(g)
    torch.cuda.synchronize()
    for g in graphs:  # warm replay
        g.replay()
    torch.cuda.synchronize()
    for ag in agents_b:
        _, dropped = ag.status()
        assert not dropped, "batch output capacity too small"
        ag.stats(reset=True)
    # Timed region: K steps of all batches.  HIP events: one pair brackets every graph replay on its stream
    # (a replay is the counter-reset kernel and the descent kernel; with several streams these spans overlap),
    # and one pair on the launch stream brackets all K steps, the other streams waiting for the first event and
    # the launch stream for their last work: the device time of the K steps.
    evs_b = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in graphs]
             for _ in range(args.steps)]
    span0, span1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    span0.record(stream)
    for sx in streams_b[1:]:
        sx.wait_event(span0)
    for step_evs in evs_b:
        for bi, (g, (e0, e1)) in enumerate(zip(graphs, step_evs)):
            sx = streams_b[bi % NS]
            with torch.cuda.stream(sx):
                e0.record(sx)
                g.replay()
                e1.record(sx)
    for sx in streams_b[1:]:
        stream.wait_stream(sx)
    span1.record(stream)
    torch.cuda.synchronize()
    wall_b = (time.perf_counter() - t0) / args.steps
    device_b = span0.elapsed_time(span1) / 1e3 / args.steps
    kern_b = sum(e0.elapsed_time(e1) for step_evs in evs_b for (e0, e1) in step_evs) / 1e3 / args.steps
    lookups_b = bytes_b = 0
    for ag in agents_b:
        st = ag.stats(reset=True)
        lookups_b += st["lookups"] / args.steps
        bytes_b += st["row_bytes"] / args.steps
    log(f"(b) {nb} log-normal queries ({int(ob_h[-1]) / nb:.0f} values/query), {nbatches} graphs on {NS} streams: "
        f"{wall_b * 1e3:.1f} ms/step (device {device_b * 1e3:.1f} ms; replays {kern_b * 1e3:.1f} ms summed over "
        f"overlapping spans), {lookups_b / wall_b / 1e9:.3f} G lookups/s, {bytes_b / device_b / 1e9:.0f} GB/s rows")

    # Traffic sweep (for counter runs): (a)'s launch, then one direct launch of every batch of (b), under each
    # launch shape of the tuner's space in turn; a one-element fill kernel marks the start of each shape's (a)
    # launch and of its (b) batches.
    # Each shape's (a) launch and (b) batches are timed with HIP events on the launch stream (best of
    # --sweep-rounds rounds; the markers are outside the timed spans).
    shape_sweep = None
    if args.traffic_sweep:
        from hibf_amd import autotune
        saved = {k: native.tune_get(k) for k in autotune.HIBF_SPACE}
        marker = torch.empty(1, dtype=torch.int64, device=dev)
        shape_sweep = []
        for si, knobs in enumerate(autotune._candidates(autotune.HIBF_SPACE)):
            autotune._apply(knobs)
            a_ms, b_ms = [], []
            for _ in range(max(1, args.sweep_rounds)):
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
                marker.fill_(si)
                hibf.membership_prepare(agent, nq)
                ev[0].record(stream)
                hibf.membership_launch(agent, va, oa, ta, ids, hs, hc)
                ev[1].record(stream)
                hibf.membership_prepare(agent, G)
                marker.fill_(1000 + si)  # and the start of its (b) batches
                ev[2].record(stream)
                for b in range(nbatches):
                    q0, q1 = b * G, min(nb, (b + 1) * G)
                    hibf.membership_launch(agent, vb, ob[q0:q1 + 1], tb[q0:q1], ids_b[b * cap_b:(b + 1) * cap_b],
                                           hs_b[q0:q1], hc_b[q0:q1])
                ev[3].record(stream)
                torch.cuda.synchronize()
                a_ms.append(ev[0].elapsed_time(ev[1]))
                b_ms.append(ev[2].elapsed_time(ev[3]))
            shape_sweep.append({"tuning": knobs, "a_ms": min(a_ms), "b_ms": min(b_ms)})
            log(f"traffic sweep shape {si}: {knobs} (a) {min(a_ms):.3f} ms (b) {min(b_ms):.2f} ms")
        autotune._apply(saved)
        agent.stats(reset=True)

    # (d) counting agent: uint16 bulk_count over every user bin for a prefix of (a), t = 0.7 L as in (a).
    count_leg = None
    ncq = min(args.hibf_count_queries, nq)
    if ncq > 0:
        thr_c = torch.from_numpy(ta_h[:ncq].astype(np.int64)).to(dev).view(torch.uint64)
        counts_c = torch.empty((ncq, n_ub), dtype=torch.uint16, device=dev)
        hibf.bulk_count_into(agent, va, oa[: ncq + 1], thr_c, counts_c, 2)
        torch.cuda.synchronize()
        agent.stats(reset=True)
        cev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(3)]
        for a, b in cev:
            a.record(stream)
            hibf.bulk_count_into(agent, va, oa[: ncq + 1], thr_c, counts_c, 2)
            b.record(stream)
        torch.cuda.synchronize()
        st = agent.stats(reset=True)
        c_ms = float(np.mean([a.elapsed_time(b) for a, b in cev]))
        out_bytes = ncq * n_ub * 2
        count_leg = {"queries": ncq, "threshold": "0.7 L", "kernel_ms": c_ms,
                     "lookups_per_s": st["lookups"] / 3 / (c_ms / 1e3),
                     "row_GBs": st["row_bytes"] / 3 / (c_ms / 1e3) / 1e9,
                     "counter_GBs": out_bytes / (c_ms / 1e3) / 1e9}
        log(f"(d) bulk_count {ncq} x 250 over {n_ub} user bins: {c_ms:.2f} ms, "
            f"{count_leg['lookups_per_s'] / 1e9:.3f} G lookups/s, {count_leg['counter_GBs']:.0f} GB/s of counters")

    # (e) reads in: k-mer hashes (k = 19) of random 250-base reads, as the tutorial search feeds membership_for.
    kmer_leg = None
    if args.reads > 0:
        import oracle as O
        from hibf_amd import native as N
        from hibf_amd.device import to_numpy
        rng_r = np.random.default_rng(args.seed + 21)
        nr, rl, k = args.reads, 250, 19
        seq = torch.from_numpy(np.frombuffer(b"ACGT", dtype=np.uint8)[rng_r.integers(0, 4, size=nr * rl)].copy()).to(dev)
        soff = torch.arange(0, (nr + 1) * rl, rl, dtype=torch.int64, device=dev).view(torch.uint64)
        hoff = torch.empty(nr + 1, dtype=torch.uint64, device=dev)
        hashes_r = torch.empty(nr * (rl - k + 1), dtype=torch.uint64, device=dev)
        total = ctypes.c_uint64()

        def hash_reads():
            N.check(N.lib().hibf_gpu_kmer_hash(ctypes.c_void_p(seq.data_ptr()), ctypes.c_void_p(soff.data_ptr()), nr, k,
                                               ctypes.c_void_p(hoff.data_ptr()), ctypes.c_void_p(hashes_r.data_ptr()),
                                               hashes_r.numel(), ctypes.byref(total),
                                               ctypes.c_void_p(stream.cuda_stream)))
        hash_reads()
        t1 = time.perf_counter()
        for _ in range(3):
            hash_reads()
        r_s = (time.perf_counter() - t1) / 3
        kmer_leg = {"reads": nr, "read_length": rl, "k": k, "hashes": int(total.value), "seconds": r_s,
                    "hashes_per_s": total.value / r_s, "GBs": (nr * rl + total.value * 8) / r_s / 1e9}
        log(f"(e) k-mer hashes of {nr} x {rl} bases (k = {k}): {r_s * 1e3:.2f} ms per call incl. offsets and sync, "
            f"{kmer_leg['hashes_per_s'] / 1e9:.2f} G hashes/s")

        # The same reads as minimisers: ungapped{19}, window_size{23}, seqan3's default seed.
        shape, window, mseed = (1 << k) - 1, 23, 0x8F3F73B5CF1C9ADE

        def minimise_reads():
            N.check(N.lib().hibf_gpu_minimiser_hash(ctypes.c_void_p(seq.data_ptr()), ctypes.c_void_p(soff.data_ptr()),
                                                    nr, ctypes.c_uint64(shape), window, ctypes.c_uint64(mseed),
                                                    ctypes.c_void_p(hoff.data_ptr()),
                                                    ctypes.c_void_p(hashes_r.data_ptr()), hashes_r.numel(),
                                                    ctypes.byref(total), ctypes.c_void_p(stream.cuda_stream)))
        minimise_reads()
        t1 = time.perf_counter()
        for _ in range(3):
            minimise_reads()
        m_s = (time.perf_counter() - t1) / 3
        mini = {"shape": shape, "window_size": window, "seed": mseed, "minimisers": int(total.value),
                "seconds": m_s, "kmers_per_s": nr * (rl - k + 1) / m_s, "reads_per_s": nr / m_s}
        nchk = min(nr, 256)
        h_m, o_m = to_numpy(hashes_r), to_numpy(hoff)
        bases = seq[: nchk * rl].cpu().numpy().tobytes()
        mini["parity_sample_ok"] = all(
            h_m[int(o_m[i]):int(o_m[i + 1])].tolist() == O.minimiser_hashes(bases[i * rl:(i + 1) * rl], shape, window,
                                                                             mseed).tolist()
            for i in range(nchk))
        if not args.skip_cpu:
            # Oracle restatement on one core (a per-k-mer hash loop, as the definition reads), ~1 s of reads.
            ncpu = max(1, min(nr, int(2000 * max(args.cpu_seconds, 0.1) / 10)))
            cb = seq[: ncpu * rl].cpu().numpy().tobytes()
            t1 = time.perf_counter()
            for i in range(ncpu):
                O.minimiser_hashes(cb[i * rl:(i + 1) * rl], shape, window, mseed)
            c_s = time.perf_counter() - t1
            mini["cpu_baseline"] = {"value": ncpu * (rl - k + 1) / c_s, "unit": "k-mers/s", "cores": 1,
                                    "kind": "port", "sample": f"oracle minimiser_hash of the first {ncpu} reads"}
        kmer_leg["minimisers"] = mini
        log(f"(e) minimisers of the same reads (k = {k}, window {window}): {m_s * 1e3:.2f} ms per call incl. offsets "
            f"and sync, {mini['kmers_per_s'] / 1e9:.2f} G k-mers/s, {total.value / nr:.1f} minimisers per read, "
            f"parity on {nchk} reads {mini['parity_sample_ok']}")
        del seq, soff, hoff, hashes_r

    # (c) plain form: one launch per query of (b), back to back.  The launches call the C-ABI directly, as a
    # C++ caller would, without the Python wrapper's buffer checks.
    npl = min(args.plain_queries, nb)
    hibf.membership_prepare(agent, 1)
    ids1 = torch.empty(1 << 16, dtype=torch.uint64, device=dev)
    c_launch = native.lib().hibf_gpu_hibf_membership_launch
    P = lambda t_: ctypes.c_void_p(t_.data_ptr())  # noqa: E731
    p_plain = [(P(ob[q:q + 2]), P(tb[q:q + 1]), P(hs_b[q:q + 1]), P(hc_b[q:q + 1])) for q in range(npl)]
    p_vb, p_ids1, p_stream = P(vb), P(ids1), ctypes.c_void_p(stream.cuda_stream)

    def plain_launch(q):
        o_, t_, s_, c_ = p_plain[q]
        native.check(c_launch(hibf.h, agent.h, p_vb, o_, 1, t_, p_ids1, ids1.numel(), s_, c_, p_stream))
    torch.cuda.synchronize()
    agent.stats(reset=True)
    t0 = time.perf_counter()
    for q in range(npl):
        plain_launch(q)
    torch.cuda.synchronize()
    wall_c = time.perf_counter() - t0
    st = agent.stats(reset=True)
    plain = st["lookups"] / wall_c
    lat = []
    for q in range(min(300, npl)):
        t1 = time.perf_counter()
        plain_launch(q)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    log(f"(c) plain per-query: {plain / 1e9:.4f} G lookups/s; latency p50 {np.percentile(lat, 50):.3f} ms "
        f"p99 {np.percentile(lat, 99):.3f} ms; batched/plain = {lookups_b / wall_b / plain:.1f}")

    # Parity on samples of (a) and (b) against the oracle on the downloaded arena, and the CPU baseline.
    parity = None
    cpu = None
    if args.check_queries > 0 or not args.skip_cpu:
        import oracle as O
        from hibf_amd.device import to_numpy
        arena = hibf.download_arena()
        geoms = [O.geom(b_, s_, h_) for (b_, s_, h_) in tab["geoms"]]
        ref = O.Hibf(geoms, arena, tab["word_offsets"], tab["tb_offsets"], tab["next_ibf"], tab["user_bin"], n_ub)
        rng = np.random.default_rng(args.seed + 9)
        ok = True
        # (a) outputs are from the last timed step; (b) from the last replay.
        for v_h, o_h, t_h, ids_t, hs_t, hc_t, n, per_batch in (
                (va_h, oa_h, ta_h, ids, hs, hc, nq, None), (vb_h, ob_h, tb_h, ids_b, hs_b, hc_b, nb, G)):
            ids_n, hs_n, hc_n = to_numpy(ids_t), to_numpy(hs_t), to_numpy(hc_t)
            for q in rng.choice(n, size=min(args.check_queries, n), replace=False):
                q = int(q)
                base = 0 if per_batch is None else (q // per_batch) * cap_b
                s0 = base + int(hs_n[q])
                got = ids_n[s0: s0 + int(hc_n[q])].tolist()
                want = ref.membership_for(v_h[int(o_h[q]):int(o_h[q + 1])], int(t_h[q]))
                if got != want:
                    ok = False
                    log(f"parity mismatch at query {q}: {got} vs {want}")
                    break
        if count_leg is not None and ok:
            got_c = counts_c.view(torch.int16).cpu().numpy().view(np.uint16)
            pick = rng.choice(ncq, size=min(args.check_queries, ncq), replace=False)
            for q in pick:
                q = int(q)
                want = ref.bulk_count(va_h[int(oa_h[q]):int(oa_h[q + 1])], int(ta_h[q]), 2)
                if not np.array_equal(got_c[q], want):
                    ok = False
                    log(f"bulk_count mismatch at query {q}")
                    break
            count_leg["parity_sample_ok"] = ok
        parity = ok
        log(f"parity vs oracle on {2 * args.check_queries} sampled queries: {parity}")
        if not args.skip_cpu:
            threads = max(1, min(16, os.cpu_count() or 1))

            def run_cpu(n, t):
                ref.membership_hits_batch(va_h[: n * 250], oa_h[: n + 1], ta_h[:n], t)
            simd = O.set_simd_counting(True)
            n_cpu, cpu_s = timed_sample(lambda n: run_cpu(n, threads), 200, nq, args.cpu_seconds)
            n_one, one_s = timed_sample(lambda n: run_cpu(n, 1), 50, nq, args.cpu_seconds / 4)
            O.set_simd_counting(False)
            # Lookups (a value checked against one visited IBF) of the same prefix, counted by the device.
            hibf.membership_prepare(agent, n_cpu)
            agent.stats(reset=True)
            hibf.membership_launch(agent, va, oa[: n_cpu + 1], ta[:n_cpu], ids, hs[:n_cpu], hc[:n_cpu])
            torch.cuda.synchronize()
            lk_cpu = agent.stats(reset=True)["lookups"]
            hibf.membership_prepare(agent, n_one)
            hibf.membership_launch(agent, va, oa[: n_one + 1], ta[:n_one], ids, hs[:n_one], hc[:n_one])
            torch.cuda.synchronize()
            lk_one = agent.stats(reset=True)["lookups"]
            cpu = {"value": lk_cpu / cpu_s, "unit": "lookups/s", "cores": threads, "kind": "port",
                   "sample": f"first {n_cpu} queries of (a) ({lk_cpu} lookups), oracle membership_for with "
                             f"{threads} OpenMP threads, one agent per thread, {cpu_s:.1f} s",
                   "query_values_per_s": n_cpu * 250 / cpu_s,
                   "one_thread": {"value": lk_one / one_s, "queries": n_one, "seconds": one_s}, **host_cpu(simd)}
            log(f"cpu baseline: {cpu['value'] / 1e6:.2f} M lookups/s on {threads} threads, "
                f"{cpu['one_thread']['value'] / 1e6:.2f} M lookups/s on 1 thread")
        del arena, ref

    # Serving leg (the serving loop on the same resident HIBF): p50/p99 from arrival to ids on the host.
    serving = None
    if not args.no_serve_leg:
        rates = [float(r) for r in args.leg_serve_rates.split(",")]
        rows = serve_sweep(hibf, tab, sizes, spec, args, rates, [args.leg_serve_window], args.leg_serve_seconds)
        serving = {"window_us": args.leg_serve_window, "seconds_of_arrivals": args.leg_serve_seconds,
                   "streams": 4, "budget_bytes": 4 << 30, "rows": rows}

    kernel_name = descent_kernel_name(native)
    per_launch_bytes = bytes_b / nbatches
    # Launches overlap on NS streams: a launch's share of the device time is the device time of a step over its
    # launches (each replay's own span, start to end, is reported beside it).
    launch_s = device_b / nbatches
    achieved = per_launch_bytes / launch_s / 1e9
    traffic = recorded_traffic("hibf_batched", args, layout)
    out = {
        "metric": "k-mer lookups/s (HIBF membership_for, batched hipGraph form)",
        "value": lookups_b / wall_b,
        "unit": "lookups/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": wall_b * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 hash / u16 counters",
        "data": "synthetic (seeded generator values, uniform uint64 noise)",
        "config": {"workload": f"batched form on the HIBF at scale: {n_ub} user bins (log-normal sizes, median 1e5, "
                               f"sigma 1.5, clipped [1e3, 5e7]), tmax 256, h=2, FPR 0.05/0.3, reference-default "
                               f"layout; {nb} queries, log-normal length median 300 [50, 5000], 60% from one user "
                               f"bin, t = uint16(0.7 L); persistent work-stealing descent in hipGraph batches of {G}",
                   "queries": nb, "values_per_step": int(ob_h[-1]), "graph_batch": G, "graphs": nbatches,
                   "ibfs": len(tab["geoms"]), "arena_gb": tab["n_words"] * 8 / 1e9, "depth": tab["depth"],
                   "tuning": {k: native.tune_get(k) for k in autotune.HIBF_SPACE},
                   "autotune": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel_name,
                     "kernel_ms_per_launch": launch_s * 1e3, "launches_per_step": nbatches,
                     "streams": NS, "replay_span_ms_mean": kern_b / nbatches * 1e3,
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "bytes_per_lookup": bytes_b / max(lookups_b, 1),
                     "random_line_ceiling": {
                         "lines_per_s": RANDOM_LINE_CEILING, "source": "profiles/random_row_gather_ceiling.md",
                         "frac": None if traffic is None else traffic / 128.0 / launch_s / RANDOM_LINE_CEILING}},
        "fixed_batch": {"queries": nq, "length": 250, "ms": wall_a * 1e3, "lookups_per_s": lookups_a / wall_a,
                        "row_gbs": bytes_a / kern_a / 1e9, "kernel_ms": kern_a * 1e3,
                        "traffic": recorded_traffic("hibf_fixed_batch", args, layout)},
        "bulk_count": count_leg,
        "kmer_hash": kmer_leg,
        "plain_per_query": {"value": plain, "unit": "lookups/s", "queries": npl,
                            "latency_ms_p50": float(np.percentile(lat, 50)),
                            "latency_ms_p99": float(np.percentile(lat, 99))},
        "batched_vs_plain": lookups_b / wall_b / plain,
        "cpu_baseline": cpu,
        "serving": serving,
        "parity_sample_ok": parity,
        "build_s": build_s,
        "layout_s": layout_s,
        "layout": {"mode": "union estimates + rearrangement (reference default)",
                   "computed": layout_times or None},
        "sketches": sketch_leg,
    }
    if shape_sweep is not None:
        out["shape_sweep"] = shape_sweep
    if not args.no_flat_leg:
        del graphs, ids_b, hs_b, hc_b, vb, ob, tb, va, oa, ta, ids, hs, hc
        torch.cuda.empty_cache()
        f = flat_measure(args, leg=True)
        out["flat_ibf"] = {k: f[k] for k in ("value", "unit", "ms_per_step", "config", "roofline", "cpu_baseline",
                                              "plain_per_query", "batched_vs_plain", "bulk_contains",
                                              "membership_for", "parity_sample_ok", "emplace_columns_ok",
                                              "build_emplace_s")}
    print(json.dumps(out), flush=True)


def descent_kernel_name(native) -> str:
    """The membership descent kernel the current launch-shape knobs select for
    an HIBF whose IBFs all use two hash functions and rows of at most 4 words."""
    if native.tune_get("hibf.wave_kernel"):
        return "hibf_wave_membership_h2_w4_kernel (+ hibf_descent_* for queries beyond its LDS id buffer)"
    nt = "true" if native.tune_get("hibf.nt_loads") else "false"
    if native.tune_get("hibf.row_per_lane"):
        return f"hibf_descent_membership_h2_row_per_lane_kernel<{nt}>"
    sliced = native.tune_get("hibf.sliced_counts")
    unroll = native.tune_get("hibf.unroll") if sliced else 2
    return f"hibf_descent_membership_h2_w4_kernel<{'true' if sliced else 'false'}, {unroll}, {nt}>"


def serve_sweep(hibf, tab, sizes, spec, args, rates, windows, seconds):
    """Serving loop on a resident HIBF for each offered rate and window:
    Poisson arrivals, log-normal lengths (median 300, at most 5,000), 60 % of
    the values from one user bin, t = 0.7 L, 4 GiB admission budget, 4
    streams; results of sampled requests checked against the oracle.
    Returns the rows of the sweep."""
    import torch
    from hibf_amd import workloads as W
    n_ub = int(sizes.size)
    rows = []
    ref = None
    for rate in rates:
        nq = int(rate * seconds)
        lengths = W.query_lengths(args.seed + 11, nq)
        # The values, their pinned copy and the generator's temporaries: about 40 bytes per value at the peak.
        need = int(lengths.sum()) * 40
        if need > args.serve_host_gb * 1e9:
            log(f"serve rate {rate:.0f} q/s skipped: its {int(lengths.sum()):.3e} query values would need about "
                f"{need / 1e9:.0f} GB of host memory (--serve-host-gb {args.serve_host_gb:.0f}; lower --serve-seconds)")
            continue
        v_h, o_h, _ = W.hibf_queries(args.seed + 12, sizes, spec["build_seed"], lengths, 0.6)
        t_h = np.minimum((0.7 * np.diff(o_h.astype(np.int64))).astype(np.int64), 65535).astype(np.uint16)
        rng = np.random.default_rng(args.seed + 13)
        arrival = np.cumsum(rng.exponential(1.0 / rate, size=nq))
        arrival -= arrival[0]
        v_in = v_h
        if not args.serve_pageable:
            # Requests land in a pinned receive buffer, as a server's network path would place them.
            v_in = torch.empty(v_h.size, dtype=torch.int64, pin_memory=True)
            v_in.numpy()[:] = v_h.view(np.int64)
        lat = ids = cnt = None
        rate_rows = []
        for win in windows:
            lat, ids, cnt, st = hibf.serve(v_in, o_h, t_h, arrival, window_us=win, streams=4,
                                           budget_bytes=4 << 30, results_per_query=16)
            row = {"rate_qps": rate, "window_us": win, "queries": nq,
                   "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                   "sustained_qps": nq / st["seconds"], "sustained_lookups_per_s": st["lookups"] / st["seconds"],
                   "batches": st["batches"], "mean_batch": nq / max(st["batches"], 1),
                   "max_in_flight_mb": st["max_in_flight_bytes"] / 2**20, "overflowed": st["overflowed"],
                   "longest_pass_ms": st["longest_pass_s"] * 1e3, "longest_launch_ms": st["longest_launch_s"] * 1e3}
            log(f"serve rate {rate:.0f} q/s window {win:.0f} us: p50 {row['p50_ms']:.3f} ms p99 {row['p99_ms']:.3f} ms, "
                f"{row['sustained_qps']:.0f} q/s, {row['sustained_lookups_per_s'] / 1e9:.2f} G lookups/s, "
                f"{st['batches']} batches (mean {row['mean_batch']:.1f}); longest loop pass "
                f"{row['longest_pass_ms']:.2f} ms, longest batch launch {row['longest_launch_ms']:.2f} ms")
            rate_rows.append(row)
        if args.check_queries > 0:
            import oracle as O
            if ref is None:
                arena = hibf.download_arena()
                geoms = [O.geom(b_, s_, h_) for (b_, s_, h_) in tab["geoms"]]
                ref = O.Hibf(geoms, arena, tab["word_offsets"], tab["tb_offsets"], tab["next_ibf"],
                             tab["user_bin"], n_ub)
            ok = True
            for q in rng.choice(nq, size=min(args.check_queries, nq), replace=False):
                q = int(q)
                want = ref.membership_for(v_h[int(o_h[q]):int(o_h[q + 1])], int(t_h[q]))
                got = ids[q, :min(int(cnt[q]), 16)].tolist()
                if not (int(cnt[q]) == len(want) and got == want[:16]):
                    ok = False
                    log(f"serving result differs at query {q}")
                    break
            for row in rate_rows:  # the last window's results were checked
                row["parity_sample_ok"] = ok if row is rate_rows[-1] else None
            log(f"serving parity vs oracle on {min(args.check_queries, nq)} sampled queries: {ok}")
        rows.extend(rate_rows)
        del v_h, v_in, o_h, t_h, arrival, lat, ids, cnt
    return rows


def run_serve(args):
    """Serving loop on the 32,768-user-bin HIBF: Poisson arrivals, log-normal
    lengths (median 300, at most 5,000), 60% from one user bin, t = 0.7 L,
    4 GiB admission budget, batching windows of 0.2-2 ms, 4 streams."""
    import torch
    from hibf_amd import workloads as W
    from hibf_amd.device import DeviceHibf

    from hibf_amd import native
    for kv in args.tune:
        k, v = kv.split("=")
        native.tune_set(k, int(v))
    torch.cuda.set_device(0)
    spec = W.HIBF_AT_SCALE
    n_ub = args.hibf_user_bins
    sizes = W.hibf_user_bin_sizes(spec["size_seed"], n_ub)
    _, counts, regs = hibf_sketch_leg(sizes, spec, None)
    layout = W.cached_layout(counts, spec["config"], registers=regs)
    del regs
    hibf = DeviceHibf.build_synthetic(sizes, spec["build_seed"], layout_text=layout, **spec["config"])
    tab = hibf.tables()
    rates = [float(r) for r in args.serve_rates.split(",")]
    windows = [float(w) for w in args.serve_windows.split(",")]
    rows = serve_sweep(hibf, tab, sizes, spec, args, rates, windows, args.serve_seconds)
    best = max(rows, key=lambda r: r["sustained_lookups_per_s"])
    out = {"metric": "p99 query latency (HIBF serving loop)", "value": best["p99_ms"], "unit": "ms",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic", "dtype": "u64 hash / u16 counters",
           "config": {"workload": f"serving on the {n_ub}-user-bin HIBF, Poisson arrivals, log-normal lengths "
                                  f"(median 300, max 5000), 60% from one user bin, t = 0.7 L, 4 GiB budget, 4 streams",
                      "query_values": "pageable, staged" if args.serve_pageable else "pinned receive buffer"},
           "sweep": rows}
    print(json.dumps(out), flush=True)


def run_serve_reads(args):
    """Serving loop for reads (the tutorial search online): user bins are the
    19-mer sets of random genomes; requests are 250-base reads sampled from
    them (1 % base errors), Poisson arrivals; bases go to the device, which
    hashes them and uses threshold uint16_t(0.7 x #k-mers).  With
    --serve-minimisers both the genomes and the reads are hashed to minimisers
    (ungapped{19}, window_size{23}, seqan3's default seed) instead."""
    import torch
    from hibf_amd import native
    from hibf_amd.device import MINIMISER_SEED, DeviceHibf, kmer_hash, minimiser_hash, to_numpy

    for kv in args.tune:
        k_, v_ = kv.split("=")
        native.tune_set(k_, int(v_))
    torch.cuda.set_device(0)
    k, read_len = 19, 250
    rng = np.random.default_rng(args.seed + 41)
    n_g, g_len = args.genomes, args.genome_length
    genome = rng.integers(0, 4, size=n_g * g_len, dtype=np.uint8)
    acgt = np.frombuffer(b"ACGT", dtype=np.uint8)
    genome_chars = acgt[genome]
    t0 = time.perf_counter()
    offs = np.arange(n_g + 1, dtype=np.uint64) * np.uint64(g_len)
    mini = (2**k - 1, 23, MINIMISER_SEED) if args.serve_minimisers else None
    hashes, hoff = minimiser_hash((genome_chars, offs), *mini) if mini else kmer_hash((genome_chars, offs), k)
    h, ho = to_numpy(hashes), to_numpy(hoff)
    ubs = [h[int(ho[i]):int(ho[i + 1])] for i in range(n_g)]
    hibf = DeviceHibf.build(ubs, tmax=64)
    tab = hibf.tables()
    log(f"serve_reads: {n_g} genomes x {g_len} bases, {int(ho[-1])} {'minimisers' if mini else 'k-mers'}, "
        f"{len(tab['geoms'])} IBFs, "
        f"built in {time.perf_counter() - t0:.1f} s")
    rates = [float(r) for r in args.serve_rates.split(",")]
    windows = [float(w) for w in args.serve_windows.split(",")]
    rows = []
    ref = None
    for rate in rates:
        nq = int(rate * args.serve_seconds)
        # Reads are generated a million at a time: the gather indices of all of
        # them at once would take 2 KB of host memory per read.
        data = np.empty(nq * read_len, dtype=np.uint8)
        for c0 in range(0, nq, 1 << 20):
            c1 = min(nq, c0 + (1 << 20))
            g = rng.integers(0, n_g, size=c1 - c0)
            start = rng.integers(0, g_len - read_len, size=c1 - c0)
            idx = (g * g_len + start)[:, None] + np.arange(read_len)[None, :]
            bases = genome_chars[idx]
            err = rng.random(bases.shape) < 0.01
            bases[err] = acgt[rng.integers(0, 4, size=int(err.sum()))]
            data[c0 * read_len:c1 * read_len] = bases.reshape(-1)
            del g, start, idx, bases, err
        r_off = np.arange(nq + 1, dtype=np.uint64) * np.uint64(read_len)
        arrival = np.cumsum(rng.exponential(1.0 / rate, size=nq))
        arrival -= arrival[0]
        for win in windows:
            lat, ids, cnt, st = hibf.serve_reads((data, r_off), arrival, k=k, fraction=0.7, window_us=win,
                                                 max_batch=65536, max_batch_bases=1 << 26, budget_bytes=4 << 30,
                                                 streams=4, results_per_query=16, minimiser=mini)
            row = {"rate_qps": rate, "window_us": win, "reads": nq,
                   "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                   "sustained_qps": nq / st["seconds"], "sustained_lookups_per_s": st["lookups"] / st["seconds"],
                   "host_link_GBs": nq * read_len / st["seconds"] / 1e9, "batches": st["batches"],
                   "hit_fraction": float(np.mean(cnt > 0)), "longest_pass_ms": st["longest_pass_s"] * 1e3}
            log(f"serve reads {rate:.0f} q/s window {win:.0f} us: p50 {row['p50_ms']:.3f} ms p99 {row['p99_ms']:.3f} ms, "
                f"{row['sustained_qps']:.0f} reads/s, {row['sustained_lookups_per_s'] / 1e9:.2f} G lookups/s, "
                f"{row['host_link_GBs']:.1f} GB/s of bases, {100 * row['hit_fraction']:.1f} % reads with hits")
            rows.append(row)
        if args.check_queries > 0:
            import oracle as O
            if ref is None:
                geoms = [O.geom(b_, s_, h_) for (b_, s_, h_) in tab["geoms"]]
                ref = O.Hibf(geoms, hibf.download_arena(), tab["word_offsets"], tab["tb_offsets"], tab["next_ibf"],
                             tab["user_bin"], n_g)
            for q in rng.choice(nq, size=min(args.check_queries, nq), replace=False):
                q = int(q)
                read = data[int(r_off[q]):int(r_off[q + 1])].tobytes().decode()
                hashes_q = O.minimiser_hashes(read, *mini) if mini else O.kmer_hashes(read, k)
                want = ref.membership_for(hashes_q, O.kmer_threshold(hashes_q.size, 0.7))
                assert int(cnt[q]) == len(want) and ids[q, :min(len(want), 16)].tolist() == want[:16], \
                    f"serving result differs at read {q}"
            log(f"serving parity vs oracle on {min(args.check_queries, nq)} sampled reads: True")
    best = max(rows, key=lambda r: r["sustained_qps"])
    out = {"metric": "reads/s (HIBF serving loop, reads in)", "value": best["sustained_qps"], "unit": "reads/s",
           "n_gpus": 1, "higher_is_better": True, "data": "synthetic", "dtype": "u8 bases / u64 hash",
           "config": {"workload": f"serving reads of {read_len} bases "
                                  + ("(minimisers of ungapped 19 in windows of 23, t = 0.7 x #minimisers)" if mini
                                     else f"(k = {k}, t = 0.7 x #k-mers)") + " on an HIBF of "
                                  f"{n_g} random genomes x {g_len} bases, tmax 64, Poisson arrivals, 4 streams"},
           "sweep": rows}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
