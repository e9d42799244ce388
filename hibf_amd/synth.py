"""Seeded synthetic inputs for the benchmark workloads (numpy, host side).

synth_values(seed, bin, i) is the same stream the device-side generator
emplace uses (fmix64 of (bin << 36 | i) + seed * golden ratio), so queries can
be drawn from exactly the values a bin received.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xFF51AFD7ED558CCD)
_M2 = np.uint64(0xC4CEB9FE1A85EC53)


def synth_values(seed: int, bins, idx) -> np.ndarray:
    """Vectorised synth value stream (wrapping uint64 arithmetic)."""
    b = np.asarray(bins, dtype=np.uint64)
    i = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        k = ((b << np.uint64(36)) | i) + np.uint64(seed) * _GOLDEN
        k ^= k >> np.uint64(33)
        k *= _M1
        k ^= k >> np.uint64(33)
        k *= _M2
        k ^= k >> np.uint64(33)
    return k


def flat_ibf_queries(seed: int, nq: int, length: int, bins: int, per_bin: int, frac_in: float,
                     emplace_seed: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """nq queries of `length` values: round(frac_in*length) drawn from the set
    one random technical bin received, the rest uniform uint64.
    Returns (values, offsets, source_bin)."""
    rng = np.random.default_rng(seed)
    k_in = int(round(frac_in * length))
    src = rng.integers(0, bins, size=nq, dtype=np.uint64)
    values = np.empty((nq, length), dtype=np.uint64)
    idx = rng.integers(0, per_bin, size=(nq, k_in), dtype=np.uint64)
    values[:, :k_in] = synth_values(emplace_seed, src[:, None], idx)
    values[:, k_in:] = rng.integers(0, np.iinfo(np.uint64).max, size=(nq, length - k_in), dtype=np.uint64,
                                    endpoint=True)
    offsets = np.arange(nq + 1, dtype=np.uint64) * np.uint64(length)
    return values.reshape(-1), offsets, src
