"""Writes hash_vectors.json: known-answer vectors for the reference hash.

The expected positions come from a pure-Python big-integer restatement of
interleaved_bloom_filter.hpp:185-195 (seeded multiply, xor-shift by
countl_zero(bin_size), golden-ratio multiply, fastrange, times
technical_bins).  It shares no code with the C oracle, so the vectors pin the
oracle's 64-bit/128-bit arithmetic.  Run from the repository root:
    python tests/golden/make_hash_vectors.py
"""
import json
import os
import random

SEEDS = (13572355802537770549, 13043817825332782213, 10650232656628343401,
         16499269484942379435, 4893150838803335377)
M = (1 << 64) - 1


def countl_zero(x: int) -> int:
    return 64 - x.bit_length()


def hash_and_fit(h: int, seed: int, bin_size: int, technical_bins: int) -> int:
    h = (h * seed) & M
    h ^= h >> countl_zero(bin_size)
    h = (h * 0x9E3779B97F4A7C15) & M
    h = (h * bin_size) >> 64
    return (h * technical_bins) & M


def main() -> None:
    rng = random.Random(20240611)
    geoms = [(1, 64), (2, 64), (8, 64), (1019, 128), (1024, 64), (8192, 64), (8192, 8192),
             (1 << 23, 8192), (158_000, 64), (123_456_789, 256), ((1 << 40) + 3, 64)]
    values = [0, 1, 2, 3, 126, 712, 237, M, M - 1, 1 << 63, 0x9E3779B97F4A7C15]
    values += [rng.getrandbits(64) for _ in range(21)]
    cases = []
    for bin_size, tbs in geoms:
        for k in range(5):
            for v in values:
                cases.append([v, k, bin_size, tbs, hash_and_fit(v, SEEDS[k], bin_size, tbs)])
    out = {"about": "value, seed index, bin_size, technical_bins, hash_and_fit (bit position of the row start)",
           "cases": [[str(x) for x in c] for c in cases]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hash_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
