"""Seeded workloads of the HIBF benchmarks (host side, numpy).

HIBF_AT_SCALE is the 32,768-user-bin HIBF (tmax 256, h=2, FPR 0.05, relaxed
0.3) with log-normal user-bin sizes.

hibf_user_bin_sizes: 32,768 user bins with log-normal sizes (median 1e5,
sigma 1.5, clipped to [1e3, 5e7]).  hibf_queries: queries whose values are
mostly drawn from one user bin's synthetic set, the rest uniform noise.
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os

import numpy as np

from .synth import synth_values


HIBF_AT_SCALE = dict(n_user_bins=32768, size_seed=11, build_seed=5,
                     config=dict(tmax=256, number_of_hash_functions=2, maximum_fpr=0.05, relaxed_fpr=0.3))


def hibf_user_bin_sizes(seed: int, n: int = 32768, median: float = 1e5, sigma: float = 1.5,
                        lo: int = 1000, hi: int = 50_000_000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = rng.lognormal(np.log(median), sigma, size=n)
    return np.clip(np.round(s), lo, hi).astype(np.uint64)


def query_lengths(seed: int, nq: int, fixed: int | None = None, median: float = 300.0,
                  sigma: float = np.log(100.0) / 6.0, lo: int = 50, hi: int = 5000) -> np.ndarray:
    if fixed is not None:
        return np.full(nq, fixed, dtype=np.int64)
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(np.log(median), sigma, size=nq)), lo, hi).astype(np.int64)


def hibf_queries(seed: int, sizes: np.ndarray, build_seed: int, lengths: np.ndarray, frac_in: float):
    """Returns (values, offsets, source_user_bin) for len(lengths) queries."""
    rng = np.random.default_rng(seed)
    nq = lengths.size
    src = rng.integers(0, sizes.size, size=nq).astype(np.uint64)
    offsets = np.zeros(nq + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(lengths)
    total = int(offsets[-1])
    values = rng.integers(0, np.iinfo(np.uint64).max, size=total, dtype=np.uint64, endpoint=True)
    k_in = np.round(frac_in * lengths).astype(np.int64)
    # positions of in-bin values: the first k_in[q] values of each query
    qid = np.repeat(np.arange(nq), k_in)
    starts = offsets[:-1].astype(np.int64)
    pos = np.repeat(starts, k_in) + (np.arange(qid.size) - np.repeat(np.cumsum(k_in) - k_in, k_in))
    ub = src[qid]
    idx = (rng.random(qid.size) * sizes[ub.astype(np.int64)].astype(np.float64)).astype(np.uint64)
    values[pos] = synth_values(build_seed, ub, idx)
    return values, offsets, src


_LAYOUT_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "layouts")


def layout_key(sizes: np.ndarray, config: dict) -> str:
    h = hashlib.sha1(np.ascontiguousarray(sizes, dtype=np.uint64).tobytes())
    h.update(json.dumps(config, sort_keys=True).encode())
    return h.hexdigest()[:16]


def cached_layout(sizes: np.ndarray, config: dict, write: bool = True, registers=None, times: dict | None = None,
                  reuse: bool = True) -> str:
    """Layout text for these user-bin sizes and config.  The host DP is
    quadratic in the number of user bins (about a minute at 32,768), so a
    computed layout is stored as layouts/<key>.layout.gz and reused.  With
    `registers` (the user bins' sketches on the device) and union estimates
    enabled, the layout uses union estimates and rearrangement, as the
    reference's default; `times` then receives the computation's seconds.
    reuse=False computes the layout even when it is stored."""
    from . import native
    sketched = registers is not None and not config.get("disable_estimate_union", False)
    key = layout_key(sizes, dict(config, layout="sketched") if sketched else config)
    path = os.path.join(_LAYOUT_DIR, key + ".layout.gz")
    if reuse and os.path.exists(path):
        with gzip.open(path, "rt") as f:
            return f.read()
    c = native.validate_config(native.make_config(number_of_user_bins=int(sizes.size), **config))
    if sketched:
        import time
        parts = [0.0, 0.0]
        t0 = time.perf_counter()
        text = native.compute_layout_sketched(c, sizes, registers.data_ptr(), parts)
        if times is not None:
            times.update(total_s=time.perf_counter() - t0, union_estimation_s=parts[0], rearrangement_s=parts[1])
    else:
        text = native.compute_layout(c, sizes)
    if write:
        try:
            os.makedirs(_LAYOUT_DIR, exist_ok=True)
            with gzip.open(path + ".tmp", "wt") as f:
                f.write(text)
            os.replace(path + ".tmp", path)
        except OSError:
            pass
    return text
