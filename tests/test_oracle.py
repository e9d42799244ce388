"""Pins the oracle (CPU restatement of the reference query path) against the
reference's own known answers and against what the mathematics fixes."""
import json
import os
import random

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _values(spec):
    if "range" in spec:
        return np.arange(spec["range"][0], spec["range"][1], dtype=np.uint64)
    return np.asarray(spec["values"], dtype=np.uint64)


def _build(case):
    ibf = O.Ibf(case["bins"], case["bin_size"], case["hash_funs"])
    for e in case["emplace"]:
        ibf.emplace(_values(e), e["bin"])
    return ibf


# ---------------------------------------------------------------- hashing --

def test_hash_known_answers_match_c_oracle():
    with open(os.path.join(GOLDEN, "hash_vectors.json")) as f:
        cases = json.load(f)["cases"]
    assert len(cases) > 1000
    for v, k, bs, tb, want in cases:
        v, k, bs, tb, want = int(v), int(k), int(bs), int(tb), int(want)
        shift = 64 - bs.bit_length()
        assert O.hash_and_fit(v, O.HASH_SEEDS[k], shift, bs, tb) == want


def test_hash_c_matches_python_bigint_random():
    rng = random.Random(7)
    for _ in range(3000):
        bs = rng.choice([1, 2, 3, 1019, 8192, 1 << 23, rng.getrandbits(40) + 1, rng.getrandbits(63) | 1])
        tb = 64 * rng.randint(1, 200)
        v = rng.getrandbits(64)
        k = rng.randrange(5)
        shift = 64 - bs.bit_length()
        assert O.hash_and_fit(v, O.HASH_SEEDS[k], shift, bs, tb) == O.hash_and_fit_py(v, O.HASH_SEEDS[k], shift, bs, tb)


def test_hash_closed_form_properties():
    rng = random.Random(11)
    for k in range(5):
        for e in range(0, 40):
            bs = 1 << e
            g = O.geom(64, bs, 5)
            assert g.hash_shift == 63 - e  # countl_zero(2^e)
            for _ in range(50):
                v = rng.getrandbits(64)
                pos = O.hash_and_fit(v, O.HASH_SEEDS[k], g.hash_shift, bs, g.technical_bins)
                assert pos % g.technical_bins == 0          # a row start
                assert pos // g.technical_bins < bs          # fastrange stays in range
    # bin_size 1: every value maps to row 0
    g = O.geom(10, 1, 2)
    assert all(O.hash_and_fit(rng.getrandbits(64), s, g.hash_shift, 1, g.technical_bins) == 0 for s in O.HASH_SEEDS)


def test_geometry_and_errors():
    g = O.geom(73, 1019, 3)  # interleaved_bloom_filter_test.cpp:134-143
    assert (g.bins, g.bin_size, g.hash_funs, g.technical_bins) == (73, 1019, 3, 128)
    assert g.technical_bins * g.bin_size == 130_432
    g = O.geom(64, 1024, 2)
    assert g.technical_bins * g.bin_size == 65_536
    with pytest.raises(ValueError, match="bins must be > 0"):
        O.Ibf(0, 32, 2)
    with pytest.raises(ValueError, match="hash functions"):
        O.Ibf(64, 32, 0)
    with pytest.raises(ValueError, match="hash functions"):
        O.Ibf(64, 32, 6)
    with pytest.raises(ValueError, match="size of a bin"):
        O.Ibf(64, 0, 2)


# ----------------------------------------------------------- IBF goldens --

def test_snippet_counting_agent(goldens):
    case = goldens["ibf_counting_agent"]
    ibf = _build(case)
    for q in case["queries"]:
        assert ibf.bulk_count(_values(q), q["width"]).tolist() == q["expected"]


def test_snippet_clear(goldens):
    case = goldens["ibf_clear"]
    ibf = _build(case)
    for step in case["steps"]:
        for b in step["clear"]:
            ibf.clear(b)
        assert ibf.bulk_count(_values(step)).tolist() == step["expected"]


def test_snippet_containment(goldens):
    case = goldens["ibf_containment"]
    assert _build(case).contains_bits(case["value"]) == case["expected"]


def test_snippet_counting_vector(goldens):
    case = goldens["ibf_counting_vector"]
    ibf = _build(case)
    counts = np.zeros(12, dtype=np.uint16)
    for v, want in zip(case["adds"], case["expected"]):
        O.counting_vector_add(counts, ibf.bulk_contains(v), ibf.g.bins)
        assert counts.tolist() == want
    assert (counts + counts).tolist() == case["doubled"]


def test_snippet_membership(goldens):
    case = goldens["ibf_membership"]
    ibf = _build(case)
    assert ibf.membership_for(case["query"], case["threshold"]) == case["expected"]


def test_insert_iterator_counts(goldens):
    case = goldens["insert_iterator_ibf"]
    ibf = _build(case)
    assert ibf.bulk_count(np.arange(*case["query_range"], dtype=np.uint64), 1).tolist() == case["expected"]


# ----------------------------------------------- IBF unit-test invariants --

def test_ibf_unit_expectations():
    # interleaved_bloom_filter_test.cpp:145-207: empty IBF -> no bits; full emplace -> all bits
    ibf = O.Ibf(64, 1024, 2)
    assert all(not any(ibf.contains_bits(v)) for v in range(64))
    for b in range(64):
        ibf.emplace(np.arange(64, dtype=np.uint64), b)
    assert all(all(ibf.contains_bits(v)) for v in range(64))
    ibf.clear(17)  # :233-259
    assert all(ibf.contains_bits(v) == [1] * 17 + [0] + [1] * 46 for v in range(64))
    # :296-331 counting: 128 bins x 128 values
    ibf = O.Ibf(128, 1024, 2)
    for b in range(128):
        ibf.emplace(np.arange(128, dtype=np.uint64), b)
    assert ibf.bulk_count(np.arange(128, dtype=np.uint64), 8).tolist() == [128] * 128
    # :356-407: only bins 63 and 126 of 127, size not divisible by 64
    ibf = O.Ibf(127, 1024, 2)
    for b in (63, 126):
        ibf.emplace(np.arange(128, dtype=np.uint64), b)
    want = [0] * 127
    want[63] = want[126] = 128
    assert ibf.bulk_count(np.arange(128, dtype=np.uint64)).tolist() == want
    assert ibf.bulk_count(np.arange(128, dtype=np.uint64), 8).tolist() == want


@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_counting_vector_wraps(width):
    # counting_vector_test.cpp:90-107 / :185-209: all bits set on max-1 -> max, then wrap to 0
    nbits = 2033
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    mx = np.iinfo(dt).max
    counts = np.full(nbits, mx - 1, dtype=dt)
    bits = np.full((nbits + 63) // 64, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    O.counting_vector_add(counts, bits, nbits)
    assert (counts == mx).all()
    O.counting_vector_add(counts, bits, nbits)
    assert (counts == 0).all()
    # every other bit
    counts[:] = 0
    bits[:] = np.uint64(0xAAAAAAAAAAAAAAAA)
    O.counting_vector_add(counts, bits, nbits)
    assert counts.tolist() == [i % 2 for i in range(nbits)]


def test_bulk_count_wraps_uint8():
    ibf = O.Ibf(3, 4096, 2)
    ibf.emplace([42], 1)
    c = ibf.bulk_count(np.full(300, 42, dtype=np.uint64), 1)
    assert c.tolist() == [0, 300 % 256, 0]
    assert ibf.bulk_count(np.full(300, 42, dtype=np.uint64), 2).tolist() == [0, 300, 0]


def _textbook_counts(bins, bin_size, h, inserts, query, width):
    """Independent Bloom-filter restatement: per-bin sets of set positions."""
    tbs = 64 * ((bins + 63) // 64)
    shift = 64 - bin_size.bit_length()
    sets = [set() for _ in range(bins)]
    for v, b in inserts:
        for k in range(h):
            sets[b].add(O.hash_and_fit_py(v, O.HASH_SEEDS[k], shift, bin_size, tbs) // tbs)
    out = []
    for b in range(bins):
        c = sum(all(O.hash_and_fit_py(v, O.HASH_SEEDS[k], shift, bin_size, tbs) // tbs in sets[b] for k in range(h))
                for v in query)
        out.append(c % (1 << (8 * width)))
    return out


@pytest.mark.parametrize("bins,bin_size,h", [(5, 97, 1), (64, 256, 2), (70, 33, 3), (130, 64, 5)])
def test_counts_match_textbook_bloom_filter(bins, bin_size, h):
    rng = random.Random(bins * 1000 + h)
    inserts = [(rng.getrandbits(64), rng.randrange(bins)) for _ in range(300)]
    query = [v for v, _ in inserts[:60]] + [rng.getrandbits(64) for _ in range(40)]
    ibf = O.Ibf(bins, bin_size, h)
    ibf.emplace([v for v, _ in inserts], [b for _, b in inserts])
    assert ibf.bulk_count(query, 2).tolist() == _textbook_counts(bins, bin_size, h, inserts, query, 2)


def test_emplace_sets_h_bits_in_its_column_only():
    rng = np.random.default_rng(3)
    for bins, bs, h in [(1, 64, 1), (65, 129, 3), (200, 50, 5)]:
        for _ in range(20):
            v = int(rng.integers(0, 2**63))
            b = int(rng.integers(0, bins))
            ibf = O.Ibf(bins, bs, h)
            ibf.emplace([v], b)
            words = ibf.words
            nz = np.flatnonzero(words)
            total = sum(bin(int(w)).count("1") for w in words[nz])
            assert 1 <= total <= h
            for wi in nz:
                w = int(words[wi])
                for bit in range(64):
                    if (w >> bit) & 1:
                        assert (wi % ibf.g.bin_words) * 64 + bit == b


def _first_toucher_occupancy(g, words0, values, bins, n_bins):
    """Occupancy increments of emplacing the pairs in order, derived another
    way than the oracle's sequential loop: a value counts iff it is the first
    pair (in order) to touch one of its bit positions and that position was
    unset before the call.  Positions from the oracle's hash (pinned above),
    first touches from numpy's stable sort."""
    h = g.hash_funs
    pos = np.array([[O.hash_and_fit(int(v), O.HASH_SEEDS[k], g.hash_shift, g.bin_size, g.technical_bins) + int(b)
                     for k in range(h)] for v, b in zip(values, bins)], dtype=np.uint64).reshape(-1)
    owner = np.repeat(np.arange(len(values)), h)
    valid = np.repeat(np.asarray(bins) < n_bins, h)
    pos, owner = pos[valid], owner[valid]
    order = np.argsort(pos, kind="stable")
    first = np.ones(pos.size, dtype=bool)
    first[1:] = pos[order][1:] != pos[order][:-1]
    fp, fo = pos[order][first], owner[order][first]
    unset = ((words0[(fp >> np.uint64(6)).astype(np.int64)] >> (fp & np.uint64(63))) & np.uint64(1)) == 0
    counted = np.zeros(len(values), dtype=bool)
    counted[fo[unset]] = True
    occ = np.zeros(g.technical_bins, dtype=np.uint64)
    np.add.at(occ, np.asarray(bins, dtype=np.int64)[counted], 1)
    return occ


def test_emplace_tracked_occupancy():
    """emplace with track_occupancy (src/interleaved_bloom_filter.cpp:119-140)."""
    # interleaved_bloom_filter_test.cpp:209-239: 64 hashes into each of the first 64 of 128 bins.
    ibf = O.Ibf(128, 512, 2)
    occ = np.zeros(128, dtype=np.uint64)
    for b in range(64):
        ibf.emplace_tracked(np.arange(64, dtype=np.uint64), b, occ)
    assert all(occ[:64] != 0) and all(occ[64:] == 0)
    for v in range(64):
        assert ibf.bulk_contains(v)[0] & ((1 << 64) - 1) == (1 << 64) - 1
    # bin_size 1: every hash of every value is the one bit of its column, so only the first value counts.
    one = O.Ibf(70, 1, 3)
    occ = np.zeros(one.g.technical_bins, dtype=np.uint64)
    one.emplace_tracked(np.arange(1, 40, dtype=np.uint64), [b % 7 for b in range(39)], occ)
    assert occ.tolist()[:70] == [1] * 7 + [0] * 63
    # bin_size 2, one hash: a bin's occupancy is the number of its two rows hit.
    two = O.Ibf(64, 2, 1)
    occ = np.zeros(64, dtype=np.uint64)
    vals = np.arange(100, 160, dtype=np.uint64)
    two.emplace_tracked(vals, np.arange(60) % 5, occ)
    for b in range(5):
        rows = {O.hash_and_fit(int(v), O.HASH_SEEDS[0], two.g.hash_shift, 2, 64) // 64 for v in vals[b::5]}
        assert occ[b] == len(rows)
    # Emplacing the same pairs again finds every bit set; out-of-range bins are skipped.
    rng = np.random.default_rng(17)
    for bins, bs, h in [(64, 97, 2), (130, 61, 3), (5, 4096, 5)]:
        ibf = O.Ibf(bins, bs, h)
        ibf.emplace(rng.integers(0, 2**63, size=50, dtype=np.int64).astype(np.uint64), 0)  # bits set before the call
        vals = rng.integers(0, 2**63, size=600, dtype=np.int64).astype(np.uint64)
        vals[300:400] = vals[:100]  # repeated values
        bs_ = rng.integers(0, bins + 2, size=vals.size).astype(np.uint64)
        occ = np.zeros(ibf.g.technical_bins, dtype=np.uint64)
        words0 = ibf.words.copy()
        ibf.emplace_tracked(vals, bs_, occ)
        assert np.array_equal(occ, _first_toucher_occupancy(ibf.g, words0, vals, bs_, bins))
        plain = O.Ibf.from_words(bins, bs, h, words0.copy())
        keep = bs_ < bins
        plain.emplace(vals[keep], bs_[keep])
        assert np.array_equal(plain.words, ibf.words)
        again = occ.copy()
        ibf.emplace_tracked(vals, bs_, again)
        assert np.array_equal(again, occ)
        for b in range(bins):  # each counted value set at most h of its column's bits
            col = sum(bin(int(w)).count("1") for w in ibf.words[b // 64::ibf.g.bin_words] & np.uint64(1 << (b % 64)))
            assert occ[b] <= np.count_nonzero(bs_ == b) and col <= h * occ[b] + sum(
                bin(int(w)).count("1") for w in words0[b // 64::ibf.g.bin_words] & np.uint64(1 << (b % 64)))


def test_simd_counting_equals_scalar():
    """The AVX-512 restatement of counting_vector += for uint16_t counters
    (counting_vector.hpp:252-278) counts exactly as the scalar path
    (:279-311): bin counts that are and are not multiples of 32, wrap-around
    past 65,535, and HIBF membership and counts built on it."""
    if not O.set_simd_counting(True):
        O.set_simd_counting(False)
        pytest.skip("this CPU has no AVX-512BW")
    try:
        rng = np.random.default_rng(23)
        cases = []
        for bins, bs, h in [(1, 64, 2), (31, 300, 2), (32, 300, 3), (33, 300, 2), (100, 500, 2), (256, 997, 2),
                            (1000, 211, 3)]:
            ibf = O.Ibf(bins, bs, h)
            vals = rng.integers(0, 2**63, size=3000, dtype=np.int64).astype(np.uint64)
            ibf.emplace(vals, rng.integers(0, bins, size=vals.size).astype(np.uint64))
            qs = [np.concatenate([vals[rng.integers(0, vals.size, 80)], rng.integers(0, 2**63, 40, dtype=np.int64)
                                  .astype(np.uint64)]) for _ in range(20)]
            qs.append(np.tile(vals[:7], 10000))  # 70,000 values: uint16 counts wrap
            cases.append((ibf, qs))
        simd = [[ibf.bulk_count(q, 2) for q in qs] for ibf, qs in cases]
        O.set_simd_counting(False)
        scalar = [[ibf.bulk_count(q, 2) for q in qs] for ibf, qs in cases]
        for a, b in zip(simd, scalar):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        assert int(simd[0][-1][0]) == 70000 % 65536  # one bin holding every value of the long query
        ibf, qs = cases[4]
        O.set_simd_counting(True)
        m_simd = [ibf.membership_for(q, 60) for q in qs]
        O.set_simd_counting(False)
        assert m_simd == [ibf.membership_for(q, 60) for q in qs]
    finally:
        O.set_simd_counting(False)


def test_synth_stream_and_parallel_emplace_forms():
    """The synthetic test-input stream is the same in C and numpy, and the
    oracle's parallel emplace forms equal its sequential emplace."""
    from hibf_amd.synth import synth_values
    rng = np.random.default_rng(5)
    for _ in range(200):
        seed, b, i = (int(x) for x in rng.integers(0, 2**20, size=3))
        assert O.synth_value(seed, b, i) == int(synth_values(seed, b, i))
    bins, bin_size, h = 200, 3001, 3
    seq = O.Ibf(bins, bin_size, h)
    for b in range(5, 190):
        seq.emplace(synth_values(9, b, np.arange(300, dtype=np.uint64)), b)
    gen = O.Ibf(bins, bin_size, h)
    gen.emplace_synthetic(9, 5, 185, 300, threads=4)
    assert np.array_equal(gen.words, seq.words)
    items = O.Ibf(bins, bin_size, h)
    # the same sets as ragged items in shuffled order, split into two ranges each
    order = rng.permutation(np.arange(5, 190))
    sets = np.repeat(order, 2)
    begins = np.tile([0, 120], order.size)
    counts = np.tile([120, 180], order.size)
    items.emplace_synthetic_items(9, sets, begins, counts, sets, threads=3)
    assert np.array_equal(items.words, seq.words)
    vals = rng.integers(0, 2**63, size=5000, dtype=np.int64).astype(np.uint64)
    bs = rng.integers(0, bins, size=vals.size).astype(np.uint64)
    a, b = O.Ibf(bins, bin_size, h), O.Ibf(bins, bin_size, h)
    a.emplace(vals, bs)
    b.emplace_parallel(vals, bs, threads=4)
    assert np.array_equal(a.words, b.words)


def test_no_false_negatives():
    rng = np.random.default_rng(5)
    ibf = O.Ibf(100, 4096, 3)
    vals = rng.integers(0, 2**63, size=2000, dtype=np.int64).astype(np.uint64)
    bins = rng.integers(0, 100, size=2000).astype(np.uint64)
    ibf.emplace(vals, bins)
    for v, b in zip(vals[:300], bins[:300]):
        assert ibf.contains_bits(int(v))[int(b)] == 1


# ------------------------------------------------------------ HIBF oracle --

def _tiny_hibf():
    """Root IBF: [UB0 split over TBs 0-1, merged->IBF1, UB3, deleted, UB4];
    IBF1: [UB1, UB2 split over TBs 1-2].  The deleted TB still holds the
    values of ub[5] (not a user bin), whose hits carry into UB4's run."""
    rng = np.random.default_rng(9)
    ub = [rng.integers(0, 2**63, size=n, dtype=np.int64).astype(np.uint64) for n in (40, 30, 30, 20, 10, 10)]
    root = O.Ibf(6, 2048, 2)
    root.emplace(ub[0][:20], 0)
    root.emplace(ub[0][20:], 1)
    root.emplace(np.concatenate([ub[1], ub[2]]), 2)
    root.emplace(ub[3], 3)
    root.emplace(ub[5], 4)
    root.emplace(ub[4], 5)
    child = O.Ibf(3, 2048, 2)
    child.emplace(ub[1], 0)
    child.emplace(ub[2][:15], 1)
    child.emplace(ub[2][15:], 2)
    M, D = O.BIN_MERGED, O.BIN_DELETED
    user_bin = [0, 0, M, 3, D, 4, 1, 2, 2]
    next_ibf = [0, 0, 1, 0, 0, 0, 1, 1, 1]
    arena = np.concatenate([root.words, child.words])
    hb = O.Hibf([root.g, child.g], arena, [0, root.words.size], [0, 6, 9], next_ibf, user_bin, 5)
    return hb, ub


def _brute_membership(ub, query, t):
    """Reference order for the tiny tree: pre-order over TBs."""
    q = set(int(x) for x in query)
    cnt = [len(q & set(int(x) for x in u)) for u in ub]
    out = []
    if cnt[0] >= t:
        out.append(0)
    if cnt[1] + cnt[2] >= t:
        if cnt[1] >= t:
            out.append(1)
        if cnt[2] >= t:
            out.append(2)
    # UB3 at TB 3; the deleted TB 4 is not reset, so its hits add to UB4's run (TB 5)
    if cnt[3] >= t:
        out.append(3)
    if cnt[4] + cnt[5] >= t:
        out.append(4)
    return out


def test_hibf_oracle_tiny_tree_bruteforce():
    hb, ub = _tiny_hibf()
    rng = np.random.default_rng(1)
    for trial in range(60):
        parts = [u[rng.integers(0, len(u), size=int(rng.integers(0, 12)))] for u in ub]
        query = np.unique(np.concatenate(parts))
        t = int(rng.integers(1, 10))
        assert hb.membership_for(query, t) == _brute_membership(ub, query, t)
        counts = hb.bulk_count(query, t, 2)
        mem = set(hb.membership_for(query, t))
        for u in range(5):
            assert (counts[u] > 0) == (u in mem)


def test_hibf_oracle_dfs_order_and_deleted_carry():
    hb, ub = _tiny_hibf()
    # values of UB1 only: merged bin (TB 2) fires, results come from the child
    assert hb.membership_for(ub[1][:8], 5) == [1]
    # UB3 then UB4 in TB order; UB0 first
    q = np.concatenate([ub[4][:6], ub[0][:6], ub[3][:6]])
    assert hb.membership_for(q, 6) == [0, 3, 4]
    # hits in the deleted TB carry into the next run (hierarchical_interleaved_bloom_filter.hpp:339-340)
    q = np.concatenate([ub[5][:4], ub[4][:2]])
    assert hb.membership_for(q, 6) == [4]
    assert hb.bulk_count(q, 6, 2).tolist() == [0, 0, 0, 0, 6]


def test_hibf_oracle_runs_cut_by_deleted_bins():
    """One IBF [UB0, deleted, UB0, deleted, UB0, UB1] with two user bins: UB0
    has three runs, so threshold 0 reports it three times (a list longer than
    number_of_user_bins), and bulk_count keeps the last run sum that reached
    the threshold (hierarchical_interleaved_bloom_filter.hpp:488-491)."""
    rng = np.random.default_rng(5)
    parts = [rng.integers(0, 2**63, size=20, dtype=np.int64).astype(np.uint64) for _ in range(6)]
    ibf = O.Ibf(6, 4096, 2)
    for b, p in enumerate(parts):
        ibf.emplace(p, b)
    D = O.BIN_DELETED
    hb = O.Hibf([ibf.g], ibf.words, [0], [0, 6], [0] * 6, [0, D, 0, D, 0, 1], 2)
    assert hb.membership_for(parts[5][:1], 0) == [0, 0, 0, 1]
    # Runs: TB 0 alone; TB 1 (deleted) carries into TB 2; TB 3 carries into TB 4.
    q = np.concatenate([parts[0][:7], parts[1][:2], parts[2][:3], parts[4][:4], parts[5][:1]])
    assert hb.membership_for(q, 5) == [0, 0]           # runs with 7 and 5 hits; 4 < 5
    assert hb.bulk_count(q, 5, 2).tolist() == [5, 0]   # the last run that reached 5
    assert hb.bulk_count(q, 4, 2).tolist() == [4, 0]   # TB 4's run (4 hits) comes last
    assert hb.bulk_count(q, 8, 2).tolist() == [0, 0]
    assert hb.membership_batch(q, np.array([0, q.size], dtype=np.uint64), [0])[0].tolist() == [0, 0, 0, 1]


@pytest.mark.parametrize("bins,bin_size,h,new_bins", [(70, 1000, 2, 130), (64, 333, 3, 65), (3, 500, 2, 64),
                                                      (130, 777, 2, 1000), (8192, 64, 3, 8193), (5, 100, 1, 5)])
def test_oracle_increase_bin_number_closed_form(bins, bin_size, h, new_bins):
    """increase_bin_number_to(n) after emplacing equals an IBF built with n
    bins and the same emplaces: rows move, a bin's bits do not."""
    rng = np.random.default_rng(bins + new_bins)
    vals = rng.integers(0, 2**63, size=3000, dtype=np.int64).astype(np.uint64)
    bidx = rng.integers(0, bins, size=vals.size).astype(np.uint64)
    grown = O.Ibf(bins, bin_size, h)
    grown.emplace(vals, bidx)
    grown.increase_bin_number_to(new_bins)
    fresh = O.Ibf(new_bins, bin_size, h)
    fresh.emplace(vals, bidx)
    assert (grown.g.bins, grown.g.bin_words, grown.g.technical_bins) == (new_bins, fresh.g.bin_words,
                                                                         fresh.g.technical_bins)
    assert np.array_equal(grown.words, fresh.words)
    with pytest.raises(ValueError):
        grown.increase_bin_number_to(new_bins - 1 if new_bins > 1 else 0)


def _kmer_brute(seq: str, k: int) -> list[int]:
    rank = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3}
    out = []
    for p in range(len(seq) - k + 1):
        h = 0
        for ch in seq[p:p + k]:
            h = h * 4 + rank.get(ch.upper(), 0)
        out.append(h)
    return out


def test_kmer_hashes_hand_examples():
    # base-4 numbers of the dna4 ranks, first letter most significant
    assert O.kmer_hashes("ACGT", 2).tolist() == [1, 6, 11]
    assert O.kmer_hashes("ACGTN", 3).tolist() == [6, 27, 44]       # N reads as A
    assert O.kmer_hashes("acgu", 4).tolist() == [0b00011011]        # lower case, U = T
    assert O.kmer_hashes("TTTT", 4).tolist() == [255]
    assert O.kmer_hashes("AC", 3).tolist() == []                   # shorter than k
    assert O.kmer_hashes("T" * 32, 32).tolist() == [2**64 - 1]     # k = 32 fills 64 bits
    assert O.kmer_threshold(30, 0.7) == 21 and O.kmer_threshold(0, 0.7) == 0
    assert O.kmer_threshold(10**6, 0.7) == 65535


@pytest.mark.parametrize("k", [1, 5, 19, 31, 32])
def test_kmer_hashes_brute_force(k):
    rng = random.Random(k)
    for _ in range(20):
        n = rng.randint(0, 80)
        seq = "".join(rng.choice("ACGTacgtNnUx") for _ in range(n))
        assert O.kmer_hashes(seq, k).tolist() == _kmer_brute(seq, k)


# ---------------------------------------------- minimisers (reads in) --
# seqan3::views::minimiser_hash (seqan3 3.4.2, the reference tutorial's
# dependency, absent here) restated in orc_minimiser_hash; pinned below by a
# second, structurally different statement of the rule, by closed forms and
# by invariants.  Parity with seqan3's own output stays unpinned.

_COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def _dna4(seq: str) -> str:
    """dna4 reading of a string: U as T, every other non-ACGT letter as A."""
    return "".join({"A": "A", "C": "C", "G": "G", "T": "T", "U": "T"}.get(ch.upper(), "A") for ch in seq)


def _canonical_values(seq: str, shape: int, seed: int) -> list[int]:
    """min(forward ^ seed, reverse ^ seed) per window, from string operations:
    the used letters of each window as a k-mer (forward), and the same shape
    laid over the reverse complement of the window (reverse)."""
    d = _dna4(seq)
    size = shape.bit_length()
    used = [t for t in range(size) if (shape >> (size - 1 - t)) & 1]
    out = []
    for i in range(len(d) - size + 1):
        win = d[i:i + size]
        rc = "".join(_COMP[ch] for ch in reversed(win))
        f = int(O.kmer_hashes("".join(win[t] for t in used), len(used))[0])
        r = int(O.kmer_hashes("".join(rc[t] for t in used), len(used))[0])
        out.append(min(f ^ seed, r ^ seed))
    return out


def _minimiser_chain(values: list[int], window: int) -> list[tuple[int, int]]:
    """The emitted (window index, position) pairs as a chain over positions:
    from the rightmost minimum of the first window, the next position is the
    first strictly smaller value within `window` positions (taken when it
    enters), else the rightmost minimum of the `window` values after the
    current one (taken when the current one leaves)."""
    n = len(values)
    if n == 0:
        return []
    w = min(window, n)

    def rightmost_min(lo, hi):
        best = lo
        for t in range(lo, hi):
            if values[t] <= values[best]:
                best = t
        return best

    p = rightmost_min(0, w)
    chain = [(0, p)]
    while True:
        smaller = next((t for t in range(p + 1, min(p + w, n - 1) + 1) if values[t] < values[p]), None)
        if smaller is not None:
            p = smaller
            chain.append((p - w + 1, p))
        elif p + w <= n - 1:
            left = p + 1  # the window in which the current position has just left
            p = rightmost_min(p + 1, p + w + 1)
            chain.append((left, p))
        else:
            return chain


def test_minimiser_hashes_equal_the_chain_form():
    rng = random.Random(11)
    for _ in range(1500):
        size = rng.randint(1, 9)
        shape = rng.randint(1 << (size - 1), (1 << size) - 1)
        window = rng.randint(size, size + 14)
        seed = rng.choice([0, O.MINIMISER_SEED, rng.getrandbits(64)])
        seq = "".join(rng.choice("ACGTACGTAAACCnU") for _ in range(rng.randint(0, 90)))
        values = _canonical_values(seq, shape, seed)
        expect = [values[p] for _, p in _minimiser_chain(values, window - size + 1)]
        assert O.minimiser_hashes(seq, shape, window, seed).tolist() == expect, (seq, shape, window, seed)


def test_minimiser_hashes_window_of_one_value_emits_every_value():
    rng = random.Random(12)
    for k in (1, 4, 19, 31, 32):
        for seed in (0, O.MINIMISER_SEED):
            seq = "".join(rng.choice("ACGT") for _ in range(200))
            fwd = O.kmer_hashes(seq, k).astype(object)
            rev = O.kmer_hashes("".join(_COMP[ch] for ch in reversed(seq)), k)[::-1].astype(object)
            expect = [min(int(f) ^ seed, int(r) ^ seed) for f, r in zip(fwd, rev)]
            assert O.minimiser_hashes(seq, (1 << k) - 1, k, seed).tolist() == expect


def test_minimiser_hashes_hand_examples():
    full = (1 << 64) - 1
    # one window of AC with k = 2, seed 0: forward AC = 1, reverse complement GT = 11
    assert O.minimiser_hashes("AC", 0b11, 2, 0).tolist() == [1]
    # palindromic k-mer ACGT: forward = reverse = 0b00011011
    assert O.minimiser_hashes("ACGT", 0b1111, 4, 0).tolist() == [27]
    # TTTT: forward 255, reverse AAAA = 0
    assert O.minimiser_hashes("TTTT", 0b1111, 4, 0).tolist() == [0]
    # seed all ones: min(~f, ~r) = ~max(f, r)
    assert O.minimiser_hashes("TTTT", 0b1111, 4, full).tolist() == [full ^ 255]
    # gapped 101 over ACG: forward A_G = 0b0010 = 2, reverse complement CGT -> C_T = 0b0111 = 7
    assert O.minimiser_hashes("ACG", 0b101, 3, 0).tolist() == [2]
    # shorter than the shape: nothing; shorter than the window: one minimum
    assert O.minimiser_hashes("ACG", 0b1111, 8, 0).tolist() == []
    values = _canonical_values("GATTACA", 0b111, 0)
    assert O.minimiser_hashes("GATTACA", 0b111, 40, 0).tolist() == [min(values)]
    with pytest.raises(ValueError):
        O.minimiser_hashes("ACGT", 0b1111, 3, 0)


def test_minimiser_hashes_invariants():
    """Every emitted value is the minimum of the window it is emitted at, every
    window's minimum is the last value emitted so far, and the values of a read
    and of its reverse complement share their window minima."""
    rng = random.Random(13)
    for _ in range(300):
        k = rng.randint(1, 12)
        window = rng.randint(k, k + 20)
        seed = rng.choice([0, O.MINIMISER_SEED])
        seq = "".join(rng.choice("ACGT") for _ in range(rng.randint(k, 150)))
        values = _canonical_values(seq, (1 << k) - 1, seed)
        w = min(window - k + 1, len(values))
        chain = _minimiser_chain(values, window - k + 1)
        got = O.minimiser_hashes(seq, (1 << k) - 1, window, seed).tolist()
        assert got == [values[p] for _, p in chain]
        minima = [min(values[j:j + w]) for j in range(len(values) - w + 1)]
        last, c = None, 0
        for j, m in enumerate(minima):
            while c < len(chain) and chain[c][0] <= j:
                last = values[chain[c][1]]
                c += 1
            assert last == m, (seq, k, window, j)
        rc = "".join(_COMP[ch] for ch in reversed(seq))
        rc_values = _canonical_values(rc, (1 << k) - 1, seed)
        assert rc_values == values[::-1]
        assert set(O.minimiser_hashes(rc, (1 << k) - 1, window, seed).tolist()) == set(minima)


# ------------------------------------------------ HyperLogLog (build side) --
# Known answers from the reference's test/unit/hibf/sketch/hyperloglog_test.cpp.

HLL_SMALL = [255881241332063154, 13132817195163223578, 5120631300412165844, 16862690419523416066,
             148518882728022940, 15892358469365346306, 10885195586503739779, 9563173945158404745]


def test_mt19937_64_standard_value():
    """The C++ standard fixes the 10,000th output of a default-seeded
    mt19937_64 ([rand.predef]); the reference's HLL tests draw from it."""
    assert int(O.mt19937_64(5489, 10000)[-1]) == 9981545732273789042


def test_hll_reference_expectations():
    # bit_widths / initialization (hyperloglog_test.cpp:31-51): an empty sketch estimates exactly 0.
    for bits in (5, 6, 12):
        regs = O.hll_sketch([], bits)
        assert regs.size == 1 << bits and not regs.any()
        assert O.hll_estimate(regs) == 0.0
    # add_and_estimate_small (:53-78): 8 values hit 7 registers; linear counting 32 ln(32/25).
    regs = O.hll_sketch(HLL_SMALL, 5)
    assert abs(O.hll_estimate(regs) - 7.899522493) < 1e-7
    assert int(np.count_nonzero(regs)) == 7
    # add_and_estimate_large (:112-128): 1,500 mt19937_64(0) values within 1.04 / 4 of the count.
    vals = O.mt19937_64(0, 1500)
    est = O.hll_estimate(O.hll_sketch(vals, 5))
    n = len(set(int(v) for v in vals))
    assert abs(est - n) <= n * 1.04 / 4
    # merge (:130-163): ten chunk sketches merged equal the full sketch.
    full = O.hll_sketch(vals, 5)
    merged = np.zeros(32, dtype=np.uint8)
    for c in range(0, 1500, 150):
        merged = O.hll_merge(merged, O.hll_sketch(vals[c:c + 150], 5))
    assert np.array_equal(merged, full) and O.hll_estimate(merged) == O.hll_estimate(full)


def test_hll_register_rule_by_brute_force():
    """Register index = top `bits` bits of the hash, rank = position of the
    first set bit after them (capped at 64 - bits + 1), max per register;
    checked with Python integers on the hash the oracle exports, and the hash
    against a big-integer restatement of the 128-bit product."""
    rng = np.random.default_rng(3)
    vals = rng.integers(0, 2**63, size=3000, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=3000).astype(np.uint64)
    for v in vals[:200]:
        p = int(v) * 0x9E3779B97F4A7C15
        assert O.hll_wyhash(int(v)) == (p & (2**64 - 1)) ^ (p >> 64)
    for bits in (5, 7, 12, 20):
        want = np.zeros(1 << bits, dtype=np.uint8)
        for v in vals:
            h = O.hll_wyhash(int(v))
            rest = format(h, "064b")[bits:]
            rank = min(rest.find("1") + 1 if "1" in rest else 65 - bits, 65 - bits)
            idx = h >> (64 - bits)
            want[idx] = max(want[idx], rank)
        assert np.array_equal(O.hll_sketch(vals, bits), want), bits


def test_hll_estimates_are_unbiased():
    """alpha for 2^bits >= 128 (0.7213 / (1 + 1.079 / m)): the mean relative
    error of 40 sketches of 20,000 distinct values is within 1 % of zero (a
    wrong alpha constant shifts every estimate by the same factor), and each
    estimate is within 4 standard errors 1.04 / sqrt(m)."""
    rng = np.random.default_rng(11)
    for bits in (8, 12):
        rel = []
        for _ in range(40):
            vals = np.unique(rng.integers(0, 2**63, size=20000, dtype=np.int64).astype(np.uint64))
            est = O.hll_estimate(O.hll_sketch(vals, bits))
            rel.append(est / vals.size - 1.0)
        se = 1.04 / np.sqrt(1 << bits)
        assert abs(np.mean(rel)) < 0.01 + 3 * se / np.sqrt(40), (bits, np.mean(rel))
        assert max(abs(r) for r in rel) < 4 * se, bits
    # The small-value branch: far fewer values than registers is linear counting, close to exact.
    vals = rng.integers(0, 2**63, size=100, dtype=np.int64).astype(np.uint64)
    assert abs(O.hll_estimate(O.hll_sketch(vals, 12)) - 100) < 3


def test_hll_union_estimates_pins():
    """Union estimates of the layout DP: the toolbox_test.cpp:52-83 known
    answers (four copies of the 5-bit sketch of 1,500 mt19937_64(0) values:
    every union estimates 1635), every (k, j) equal to the estimate of the
    register-wise max of sketches k .. j taken with numpy, and the union of a
    set with itself equal to the set's own estimate."""
    s = O.hll_sketch(O.mt19937_64(0, 1500), 5)
    assert O.hll_union_estimates(np.stack([s] * 4), [0, 1, 2, 3]).tolist() == [1635] * 6
    rng = np.random.default_rng(12)
    for bits in (5, 7, 10):
        sets = [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
                for n in rng.integers(0, 3000, size=9)]
        regs = np.stack([O.hll_sketch(v, bits) for v in sets])
        pos = rng.permutation(9)
        got = O.hll_union_estimates(regs, pos)
        for j in range(1, 9):
            for k in range(j):
                want = int(O.hll_estimate(np.maximum.reduce(regs[pos[k:j + 1]])))
                assert int(got[j * (j - 1) // 2 + k]) == want, (bits, k, j)
        twice = O.hll_union_estimates(np.stack([regs[3], regs[3]]), [0, 1])
        assert int(twice[0]) == int(O.hll_estimate(regs[3]))
