"""Autotuner of the query kernels' launch shapes.

The reference has no device tiles; here the count and descent kernels take
process-wide launch-shape knobs (hibf_gpu_tune_set).  The tuner times every
candidate shape of a search space on the caller's own resident batch, with HIP
events on the launch stream, and requires each candidate's output to equal the
default shape's output word for word: results never depend on the shape, and a
shape that changed them would be a bug, so the tuner raises instead of picking
it.  The fastest shape (the default unless another is faster by more than
MARGIN) is applied and returned in a record (knobs, timings,
device, workload) that save_record() writes as JSON and apply_record() applies
again in a later process.

Search spaces:
  flat bulk_count (wide rows, bin_words >= 64 and even):
    count.counter_bits  LDS counter width, 16-bit pairs or 8-bit quads
    count.wide_unroll   values in flight per wave (1, 2, 4, 8)
    count.wide_waves    waves per query block (1, 2, 4)
    count.wide_nt_loads non-temporal row loads (0, 1)
  HIBF membership descent (one launch at a time, tune_hibf_membership_launch, or
  batches overlapped over several streams, tune_hibf_membership_batches):
    hibf.wave_kernel    one query per wave (1) or per block (0)
    hibf.row_per_lane   rows of <= 4 words: one value and its whole row per
                        lane (1), or lanes sharing a value's row (0)
    hibf.sliced_counts  rows of <= 4 words: bit-sliced per-lane counts (1) or
                        one LDS atomic per set bit (0)
    hibf.nt_loads       non-temporal row loads (0, 1)
    hibf.block_waves    query per block: waves per block (4, 2, 8)
    hibf.unroll         rows of <= 4 words, sliced counts: value groups whose
                        rows a wave loads at once (2, 4)
    hibf.stream_kib     rows of <= 4 words: IBFs of at least this many KiB take
                        non-temporal row loads (64 MiB; 0: none)
"""
from __future__ import annotations

import itertools
import json
import time
from typing import Callable

import torch

from . import native

FLAT_SPACE = {
    "count.counter_bits": (16, 8),
    "count.wide_unroll": (1, 2, 4, 8),
    "count.wide_waves": (1, 2, 4),
    "count.wide_nt_loads": (0, 1),
}
# A candidate replaces the default shape only when it is faster by more than
# this fraction: smaller differences are within the run-to-run noise of a
# sub-batch timing and need not carry over to the full batch.
MARGIN = 0.02

HIBF_SPACE = {
    "hibf.wave_kernel": (0, 1),
    "hibf.row_per_lane": (0, 1),
    "hibf.sliced_counts": (1, 0),
    "hibf.nt_loads": (0, 1),
    "hibf.block_waves": (4, 2, 8),
    "hibf.unroll": (4, 2),
    "hibf.stream_kib": (1 << 16, 0),
}


_SIGNED = {torch.uint8: torch.uint8, torch.uint16: torch.int16, torch.uint32: torch.int32, torch.uint64: torch.int64}


def _equal(a: torch.Tensor, b: torch.Tensor) -> bool:
    """Word-for-word equality (unsigned tensors compared through signed views)."""
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    t = _SIGNED.get(a.dtype, a.dtype)
    return torch.equal(a.view(t), b.view(t))


def _candidates(space: dict) -> list[dict]:
    keys = list(space)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(space[k] for k in keys))]


def _apply(knobs: dict) -> None:
    for k, v in knobs.items():
        native.tune_set(k, int(v))


def _time_ms(fn: Callable[[], None], reps: int) -> float:
    """Mean duration of fn's launches (HIP events on the current stream)."""
    stream = torch.cuda.current_stream()
    fn()  # warm-up (also the output the caller compares)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in evs:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return float(sum(a.elapsed_time(b) for a, b in evs) / reps)


def _tune(space: dict, run: Callable[[], None], output: Callable[[], object], same: Callable[[object, object], bool],
          reps: int, workload: dict) -> dict:
    saved = {k: native.tune_get(k) for k in space}
    results = []
    reference = None
    # Warm the device (TLB, caches, clocks) on the current shape before the first
    # candidate is timed, so that the first candidate is not charged for it.
    for _ in range(3):
        run()
    torch.cuda.synchronize()
    try:
        for knobs in _candidates(space):
            _apply(knobs)
            ms = _time_ms(run, reps)
            out = output()
            if reference is None:
                reference = out
            elif not same(reference, out):
                raise RuntimeError(f"launch shape {knobs} changed the results")
            results.append({"knobs": knobs, "ms": ms})
    except BaseException:
        _apply(saved)
        raise
    best = min(results, key=lambda r: r["ms"])
    default = next((r for r in results if all(r["knobs"][k] == saved[k] for k in space)), None)
    if default is not None and best["ms"] > default["ms"] * (1.0 - MARGIN):
        best = default  # within timing noise of the default shape: keep it
    _apply(best["knobs"])
    return {
        "device": torch.cuda.get_device_name(0),
        "arch": getattr(torch.cuda.get_device_properties(0), "gcnArchName", None),
        "workload": workload,
        "best": best["knobs"],
        "best_ms": best["ms"],
        "default_ms": default["ms"] if default is not None else None,
        "margin": MARGIN,
        "candidates": results,
        "reps": reps,
        "created": time.strftime("%Y-%m-%dT%H:%M:%S"),
    }


def tune_flat_count(ibf, values: torch.Tensor, offsets: torch.Tensor, counter_bytes: int = 2, reps: int = 3,
                    space: dict | None = None) -> dict:
    """Tunes bulk_count of a DeviceIbf on a resident CSR batch (uint64 values
    and offsets on the device)."""
    g = ibf.geometry
    nq = offsets.numel() - 1
    dtype = {1: torch.uint8, 2: torch.uint16, 4: torch.uint32, 8: torch.uint64}[counter_bytes]
    out = torch.empty((nq, int(g.bins)), dtype=dtype, device=offsets.device)

    def run():
        ibf.bulk_count_into(values, offsets, out, counter_bytes)

    workload = {"kind": "flat bulk_count", "bins": int(g.bins), "bin_size": int(g.bin_size),
                "hash_functions": int(g.hash_funs), "queries": nq, "values": int(values.numel()),
                "counter_bytes": counter_bytes}
    return _tune(space or FLAT_SPACE, run, lambda: out.clone(), _equal, reps, workload)


def tune_hibf_membership(hibf, values: torch.Tensor, offsets: torch.Tensor, thresholds: torch.Tensor, reps: int = 2,
                         space: dict | None = None) -> dict:
    """Tunes the HIBF membership descent of a DeviceHibf on a resident batch
    (uint64 values/offsets, uint16 thresholds on the device).  Outputs are
    compared as per-query id lists in the reference's order."""
    state = {}

    def run():
        state["out"] = hibf.membership_for(values, offsets, thresholds)

    def output():
        ids, hoff = state["out"]
        return ids.clone(), hoff.clone()

    def same(a, b):
        return _equal(a[0], b[0]) and _equal(a[1], b[1])

    tab = hibf.tables()
    workload = {"kind": "hibf membership", "ibfs": len(tab["geoms"]), "user_bins": int(tab["n_user_bins"]),
                "queries": offsets.numel() - 1, "values": int(values.numel())}
    return _tune(space or HIBF_SPACE, run, output, same, reps, workload)


def tune_hibf_membership_launch(hibf, agent, values: torch.Tensor, offsets: torch.Tensor, thresholds: torch.Tensor,
                                ids: torch.Tensor, hit_start: torch.Tensor, hit_count: torch.Tensor, reps: int = 5,
                                space: dict | None = None) -> dict:
    """Tunes the HIBF membership descent in its asynchronous launch form
    (hibf_gpu_hibf_membership_launch, the form a hipGraph batch replays) on one
    resident batch: the same kernels, grid and batch size as the caller's
    timed launches, without the synchronous form's result download.  Outputs
    are compared as per-query id lists: where a query's list lands in `ids`
    follows the work-stealing schedule, its content and order do not."""
    nq = offsets.numel() - 1
    keys = list(space or HIBF_SPACE)
    state = {"shape": None}

    def run():
        shape = tuple(native.tune_get(k) for k in keys)
        if shape != state["shape"]:  # the per-block workspace depends on the shape
            hibf.membership_prepare(agent, nq)
            state["shape"] = shape
        hibf.membership_launch(agent, values, offsets, thresholds, ids, hit_start, hit_count)

    def output():
        start = hit_start.view(torch.int64)
        count = hit_count.view(torch.int64)
        total = int(count.sum())
        first = torch.cumsum(count, 0) - count  # each query's first position in the packed lists
        pos = torch.repeat_interleave(start - first, count) + torch.arange(total, device=count.device)
        return count.clone(), ids.view(torch.int64)[pos].clone()

    def same(a, b):
        return torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])

    tab = hibf.tables()
    workload = {"kind": "hibf membership (launch form)", "ibfs": len(tab["geoms"]),
                "user_bins": int(tab["n_user_bins"]), "queries": nq, "values": int(values.numel())}
    return _tune(space or HIBF_SPACE, run, output, same, reps, workload)


def tune_hibf_membership_batches(hibf, agents, streams, values: torch.Tensor, offsets: torch.Tensor,
                                 thresholds: torch.Tensor, ids: torch.Tensor, hit_start: torch.Tensor,
                                 hit_count: torch.Tensor, batch: int, reps: int = 3, space: dict | None = None) -> dict:
    """Tunes the HIBF membership descent in the form a caller overlaps its
    batches in: the queries of `offsets` cut into batches of `batch`, batch b
    launched (hibf_gpu_hibf_membership_launch) on streams[b % len(streams)]
    with agents[b % len(agents)], all batches timed as one span (the other
    streams wait for its start and the current stream for their ends).
    streams[0] must be the current stream.  `ids` is split evenly over the
    batches (a batch's hit_start values count from its slice).  Outputs are
    compared as per-query id lists."""
    nq = offsets.numel() - 1
    nb = (nq + batch - 1) // batch
    cap = ids.numel() // max(nb, 1)
    ns = len(streams)
    keys = list(space or HIBF_SPACE)
    state = {"shape": None}

    def run():
        shape = tuple(native.tune_get(k) for k in keys)
        if shape != state["shape"]:  # the per-block workspace depends on the shape
            for ag in agents:
                hibf.membership_prepare(ag, batch)
            state["shape"] = shape
        main = torch.cuda.current_stream()
        start = torch.cuda.Event()
        start.record(main)
        for sx in streams[1:]:
            sx.wait_event(start)
        for b in range(nb):
            q0, q1 = b * batch, min(nq, (b + 1) * batch)
            hibf.membership_launch(agents[b % len(agents)], values, offsets[q0:q1 + 1], thresholds[q0:q1],
                                   ids[b * cap:(b + 1) * cap], hit_start[q0:q1], hit_count[q0:q1],
                                   stream=streams[b % ns])
        for sx in streams[1:]:
            main.wait_stream(sx)

    def output():
        # A batch's hit_start values count from its own slice of ids.
        slice0 = torch.arange(nq, device=hit_start.device, dtype=torch.int64) // batch * cap
        start = hit_start.view(torch.int64) + slice0
        count = hit_count.view(torch.int64)
        total = int(count.sum())
        first = torch.cumsum(count, 0) - count  # each query's first position in the packed lists
        pos = torch.repeat_interleave(start - first, count) + torch.arange(total, device=count.device)
        return count.clone(), ids.view(torch.int64)[pos].clone()

    def same(a, b):
        return torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])

    tab = hibf.tables()
    workload = {"kind": f"hibf membership (launch form, {nb} batches of {batch} over {ns} streams)",
                "ibfs": len(tab["geoms"]), "user_bins": int(tab["n_user_bins"]), "queries": nq,
                "values": int(values.numel())}
    return _tune(space or HIBF_SPACE, run, output, same, reps, workload)


def save_record(record: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(record, f, indent=1)
        f.write("\n")


def load_record(path: str) -> dict:
    with open(path) as f:
        return json.load(f)


def apply_record(record: dict) -> dict:
    """Applies the best shape of a saved record; returns the knobs applied."""
    _apply(record["best"])
    return dict(record["best"])
