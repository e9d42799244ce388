"""Writes tests/golden/robin_hood_orders.json: iteration orders of the
reference's own robin_hood::unordered_flat_set<uint64_t> (the set the
reference chunks split user bins from, src/build/insert_into_ibf.cpp:28-47)
for scripts of set operations, as printed by oracle/_ref/robin_hood_ref
(built by `make -C oracle ref` from oracle/ref_robin_hood.cpp and the
reference tree's include/hibf/contrib/robin_hood.hpp).

Run from the repository root where the reference tree is present:
    make -C oracle ref && python tests/golden/make_robin_hood_orders.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "robin_hood_ref")
M64 = (1 << 64) - 1


def _colliding_keys(count, home_bits=20, seed=7):
    """Keys whose home slot is 0 in every table of up to 2^home_bits slots
    under the set's initial hash multiplier: the set's hash steps are
    bijections (xor-shift by 33, multiplication by odd constants), so each
    key is obtained by inverting them from a chosen hash value."""
    c1, mult = 0xFF51AFD7ED558CCD, 0xC4CEB9FE1A85EC53
    inv_c1, inv_mult = pow(c1, -1, 1 << 64), pow(mult, -1, 1 << 64)
    keys = []
    x = seed
    for i in range(count):
        x = (x * 6364136223846793005 + 1442695040888963407) & M64
        h3 = ((x >> (5 + home_bits)) << (5 + home_bits)) | (i & 31)  # home bits zero, info bits vary
        h2 = h3 ^ (h3 >> 33)
        h1 = (h2 * inv_mult) & M64
        t = h1 ^ (h1 >> 33)
        u = (t * inv_c1) & M64
        keys.append(u ^ (u >> 33))
    return keys


def cases():
    out = []
    # small sets, whole orders
    for n in (1, 2, 3, 5, 6, 7, 8, 9, 13, 20, 40, 100, 211):
        out.append(("small_%d" % n, f"gen 0 3 {n} 0 {n}\ndump 0 1\n"))
    # growth through many table sizes and info-width halvings
    for n in (1000, 7000, 65536, 100000, 250000, 1000000, 3000000):
        out.append(("grow_%d" % n, f"gen 0 5 {n} 0 {n}\ndump 0 0\n"))
    # reuse after clear keeps the table: a large set, then smaller and larger ones in the same set
    out.append(("reuse_after_clear",
                "gen 1 9 1 0 1000000\ndump 1 0\nclear 1\ngen 1 9 2 0 50000\ndump 1 0\nclear 1\n"
                "gen 1 9 3 0 300\ndump 1 1\nclear 1\ngen 1 9 4 0 3000000\ndump 1 0\n"))
    # a set left marked full by its last insertion: the next set in it starts with half-width info bytes
    out.append(("full_flag_carried_by_clear",
                "gen 2 11 1416 0 14304\ndump 2 0\nclear 2\ngen 2 12 0 0 20000\ndump 2 0\nclear 2\n"
                "gen 2 12 1 0 300\ndump 2 1\n"))
    # duplicates in the input
    out.append(("duplicates", "genmod 3 13 5 200000 77777\ndump 3 0\nclear 3\ngenmod 3 13 6 5000 33\ndump 3 1\n"))
    # a parent set filled from children in their iteration orders (update_parent_kmers), then reused
    out.append(("merges", "gen 4 17 1 0 120000\ngen 5 17 2 0 3000\ngen 6 17 3 0 400000\nmerge 7 4\nmerge 7 5\n"
                          "merge 7 6\ndump 7 0\nclear 7\ngen 7 17 9 0 90000\ndump 7 0\nmerge 8 5\ndump 8 0\n"))
    # keys sharing one home slot: long runs, info halvings down to width 2, then a rehash with a new hash
    # multiplier at the same table size; the multiplier stays with the set through clear()
    keys = _colliding_keys(300)
    out.append(("one_home_slot", "vals 9 " + " ".join(str(k) for k in keys) + "\ndump 9 1\nclear 9\n"
                                 "gen 9 19 1 0 200\ndump 9 1\n"))
    out.append(("extreme_values", f"vals 10 0 1 {M64} {M64 - 1} 4294967296 9223372036854775808\ndump 10 1\n"))
    # the pattern of a build's fullest top-level bin: a large union, then split user bins in the same set
    out.append(("union_then_split_bins",
                "gen 11 23 1 0 700000\ngen 12 23 2 0 900000\nmerge 13 11\nmerge 13 12\ndump 13 0\nclear 13\n"
                "gen 13 23 3 0 250000\ndump 13 0\nclear 13\ngen 13 23 4 0 1700000\ndump 13 0\nclear 13\n"
                "gen 13 23 5 0 10\ndump 13 1\n"))
    return out


def main():
    records = []
    for name, script in cases():
        r = subprocess.run([EXE], input=script, capture_output=True, text=True, check=True)
        records.append({"name": name, "script": script, "output": r.stdout})
    with open(os.path.join(HERE, "robin_hood_orders.json"), "w") as f:
        json.dump({"about": "Iteration orders of the reference's robin_hood::unordered_flat_set<uint64_t> "
                            "(robin_hood 3.11.5), printed by oracle/ref_robin_hood.cpp; written by "
                            "tests/golden/make_robin_hood_orders.py",
                   "cases": records}, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()
