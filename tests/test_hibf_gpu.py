"""GPU parity of the HIBF path: build (host layout + device bulk emplace),
membership_for in the reference's depth-first order, and bulk_count, against
the oracle and the reference's own known answers."""
import numpy as np
import pytest

import oracle as O
from hibf_helpers import expected_arena, explicit_split_chunks, oracle_of, synthetic_split_chunks

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _csr(qs):
    from hibf_amd.device import csr
    return csr(qs)


def _membership_lists(dev, qs, thresholds):
    from hibf_amd.device import to_numpy
    values, offsets = _csr(qs)
    ids, hoff = dev.membership_for(values, offsets, thresholds)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    return [ids[hoff[i]:hoff[i + 1]].tolist() for i in range(len(qs))]


# ----------------------------------------------------------- goldens ------

def test_snippet_hibf_two_user_bins(goldens):
    from hibf_amd.device import DeviceHibf
    _say("hibf snippet")
    case = goldens["hibf_two_user_bins"]
    dev = DeviceHibf.build(case["user_bins"])
    for q in case["queries"]:
        got = _membership_lists(dev, [q["values"]], [q["threshold"]])[0]
        assert got == q["expected"]
        if "expected_sorted" in q:
            assert sorted(got) == q["expected_sorted"]


def test_readme_golden(goldens):
    from hibf_amd.device import DeviceHibf
    _say("readme")
    case = goldens["readme"]
    dev = DeviceHibf.build(case["user_bins"], **case["config"])
    q1, q2 = case["queries"]
    assert _membership_lists(dev, [q1["values"]], [q1["threshold"]])[0] == q1["expected"]
    r2 = _membership_lists(dev, [list(range(*q2["range"]))], [q2["threshold"]])[0]
    assert sorted(r2) == q2["expected_sorted"]


def test_build_from_layout_text(goldens):
    from hibf_amd.device import DeviceHibf
    _say("build from layout")
    case = goldens["layout_two_user_bins"]
    lines = [f"#TOP_LEVEL_IBF fullest_technical_bin_idx:{case['top_level_max_bin_id']}",
             "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS"]
    for r in case["user_bin_records"]:
        lines.append(f"{r['idx']}\t{r['storage_TB_id']}\t{r['number_of_technical_bins']}")
    dev = DeviceHibf.build(case["user_bins"], layout_text="\n".join(lines) + "\n", **case["config"])
    t = dev.tables()
    assert t["geoms"][0][0] == 64
    assert t["user_bin"].tolist() == [1] * 34 + [0] * 30
    got = _membership_lists(dev, [case["query"]], [case["threshold"]])[0]
    assert sorted(got) == case["expected_sorted"]


# ---------------------------------- reference unit-test expectations ------

def test_three_level_hibf():
    """hierarchical_interleaved_bloom_filter_test.cpp:112-154."""
    from hibf_amd.device import DeviceHibf
    _say("three level")
    ubs = [np.arange(u * 16, u * 16 + 20, dtype=np.uint64) for u in range(4097)]
    dev = DeviceHibf.build(ubs, maximum_fpr=0.001, threads=4, tmax=64, disable_estimate_union=True,
                           disable_rearrangement=True)
    assert dev.tables()["depth"] >= 3
    qs, ts, want = [], [], []
    for ub in range(0, 4096, 100):
        s = ub * 16
        qs.append(np.arange(s + 16, s + 20, dtype=np.uint64))
        want.append([ub, ub + 1])
        qs.append(np.arange(s + 5, s + 9, dtype=np.uint64))
        want.append([ub])
    got = _membership_lists(dev, qs, 4)
    assert [sorted(g) for g in got] == want
    ref = oracle_of(dev)
    for q, g in zip(qs, got):
        assert g == ref.membership_for(q, 4)


def test_unevenly_sized_and_similar_user_bins():
    """hierarchical_interleaved_bloom_filter_test.cpp:164-231."""
    from hibf_amd.device import DeviceHibf
    _say("uneven + similar")
    opts = dict(maximum_fpr=0.001, threads=4, tmax=64, disable_estimate_union=True, disable_rearrangement=True)
    ubs = [np.arange((u + 1) ** 2, (u + 2) ** 2 - 1, dtype=np.uint64) for u in range(500)]
    dev = DeviceHibf.build(ubs, **opts)
    qs = [np.arange((u + 1) ** 2 + 5, (u + 1) ** 2 + 9, dtype=np.uint64) for u in range(5, 499, 20)]
    assert _membership_lists(dev, qs, 4) == [[u] for u in range(5, 499, 20)]
    ubs = [np.arange(u * 5, u * 5 + 35, dtype=np.uint64) for u in range(1000)]
    dev = DeviceHibf.build(ubs, **opts)
    qs = [np.arange(u * 5, u * 5 + 5, dtype=np.uint64) for u in range(6, 994, 20)]
    got = _membership_lists(dev, qs, 5)
    assert [sorted(g) for g in got] == [list(range(u - 6, u + 1)) for u in range(6, 994, 20)]


def test_counting_agent_reference_bounds():
    """hierarchical_interleaved_bloom_filter_test.cpp:233-295."""
    from hibf_amd.device import DeviceHibf, to_numpy
    _say("counting bounds")
    opts = dict(maximum_fpr=0.001, threads=4, tmax=64, disable_estimate_union=True, disable_rearrangement=True)
    dev = DeviceHibf.build([np.arange(500, dtype=np.uint64)] * 500, **opts)
    values, offsets = _csr([np.arange(500, dtype=np.uint64)])
    c = to_numpy(dev.bulk_count(values, offsets, 1, counter_bytes=4))[0]
    assert (c >= 500).all() and (c <= 502).all()
    ref = oracle_of(dev)
    assert np.array_equal(c, ref.bulk_count(np.arange(500, dtype=np.uint64), 1, 4))
    ubs = [np.arange((u + 1) ** 2, (u + 2) ** 2 - 1, dtype=np.uint64) for u in range(500)]
    dev = DeviceHibf.build(ubs, **opts)
    q = np.arange(1, 251_000, dtype=np.uint64)
    values, offsets = _csr([q])
    c = to_numpy(dev.bulk_count(values, offsets, 1, counter_bytes=4))[0]
    assert all(c[u] >= 2 * u + 2 for u in range(500))
    assert np.array_equal(c, oracle_of(dev).bulk_count(q, 1, 4))


# ------------------------------------------------- parity vs oracle -------

def _random_user_bins(rng, n, lo, hi):
    sizes = rng.integers(lo, hi, size=n)
    return [rng.integers(0, 2**63, size=int(s), dtype=np.int64).astype(np.uint64) for s in sizes]


def _mixed_queries(rng, ubs, n, length):
    qs = []
    for _ in range(n):
        k = int(rng.integers(1, 4))
        src = rng.integers(0, len(ubs), size=k)
        parts = [ubs[s][rng.integers(0, len(ubs[s]), size=int(length * 0.8) // k)] for s in src]
        noise = rng.integers(0, 2**63, size=length - sum(len(p) for p in parts), dtype=np.int64).astype(np.uint64)
        qs.append(np.concatenate(parts + [noise]))
    return qs


CASES = [  # (n_ub, size range, config)
    (2, (10, 40), dict()),
    (64, (50, 400), dict(number_of_hash_functions=2)),
    (70, (50, 400), dict(tmax=64, disable_estimate_union=True)),
    (300, (20, 300), dict(tmax=64, number_of_hash_functions=3)),
    (900, (5, 200), dict(tmax=64, maximum_fpr=0.01, number_of_hash_functions=1)),
    (5000, (2, 60), dict(tmax=64, relaxed_fpr=0.5)),
    (700, (20, 200), dict(tmax=512)),       # 8-word rows: lanes-per-value counting in the wave kernel
    (1500, (20, 200), dict(tmax=128, number_of_hash_functions=4)),
]


_DESCENT_VARIANTS = {  # (hibf.wave_kernel, hibf.row_per_lane, hibf.sliced_counts, hibf.block_waves, hibf.unroll,
    #  hibf.stream_kib)
    "wave": (1, 0, 1, 4, 4, 1 << 16),
    "block-row-per-lane": (0, 1, 1, 4, 4, 1 << 16),
    "block-lanes-per-value-sliced": (0, 0, 1, 4, 4, 1 << 16),
    "block-lanes-per-value-sliced-unroll-2": (0, 0, 1, 4, 2, 1 << 16),
    "block-lanes-per-value-atomics": (0, 0, 0, 4, 4, 1 << 16),
    "block-2-waves": (0, 0, 1, 2, 4, 1 << 16),
    "block-8-waves": (0, 0, 1, 8, 4, 1 << 16),
    "block-streamed-ibfs": (0, 0, 1, 4, 4, 1),  # every IBF of 1 KiB and more with non-temporal row loads
}


@pytest.fixture(params=list(_DESCENT_VARIANTS))
def descent_kernel(request):
    """Runs a test with each descent kernel: query per wave; query per block
    with one value and its row per lane; query per block with lanes sharing a
    value's row, counting through bit-sliced registers or with one LDS atomic
    per set bit, with two or four value groups of rows in flight per wave;
    blocks of 2 and 8 waves; and every IBF read with non-temporal row loads
    (hibf.stream_kib = 1).  Yields the variant's name."""
    from hibf_amd import native
    keys = ("hibf.wave_kernel", "hibf.row_per_lane", "hibf.sliced_counts", "hibf.block_waves", "hibf.unroll",
            "hibf.stream_kib")
    saved = {k: native.tune_get(k) for k in keys}
    for k, v in zip(keys, _DESCENT_VARIANTS[request.param]):
        native.tune_set(k, v)
    yield request.param
    for k, v in saved.items():
        native.tune_set(k, v)


@pytest.mark.parametrize("n_ub,sizes,config", CASES)
def test_hibf_parity_random(n_ub, sizes, config, descent_kernel):
    from hibf_amd.device import DeviceHibf, to_numpy
    _say(f"hibf parity n_ub={n_ub} {config} kernel={descent_kernel}")
    rng = np.random.default_rng(n_ub)
    ubs = _random_user_bins(rng, n_ub, *sizes)
    dev = DeviceHibf.build(ubs, **config)
    assert np.array_equal(dev.download_arena(), expected_arena(dev, ubs, split_chunks=explicit_split_chunks(dev, ubs))), \
        "bit matrices differ from the reference build's"
    ref = oracle_of(dev)
    qs = _mixed_queries(rng, ubs, 60, 120) + [np.zeros(0, dtype=np.uint64), ubs[0][:1]]
    thresholds = [int(t) for t in rng.integers(0, 60, size=len(qs))]
    thresholds[0], thresholds[1], thresholds[-1] = 0, 1, 0   # t = 0 reports every user bin (overflow path)
    got = _membership_lists(dev, qs, thresholds)
    for q, t, g in zip(qs, thresholds, got):
        assert g == ref.membership_for(q, t)
    values, offsets = _csr(qs)
    cthr = [max(1, t) for t in thresholds]
    for width in (1, 2, 4, 8):
        c = to_numpy(dev.bulk_count(values, offsets, cthr, counter_bytes=width))
        assert np.array_equal(c, ref.bulk_count_batch(values, offsets, cthr, width)), width


def _tiny_tree_with_deleted_bin():
    """Root IBF: [UB0 split over TBs 0-1, merged->IBF1, UB3, deleted, UB4];
    IBF1: [UB1, UB2 split over TBs 1-2].  The deleted TB 4 holds the values of
    ub[5], which is not a user bin."""
    rng = np.random.default_rng(9)
    ub = [rng.integers(0, 2**63, size=n, dtype=np.int64).astype(np.uint64) for n in (40, 30, 30, 20, 10, 10)]
    root = O.Ibf(6, 2048, 2)
    for vals, b in ((ub[0][:20], 0), (ub[0][20:], 1), (np.concatenate([ub[1], ub[2]]), 2), (ub[3], 3), (ub[5], 4),
                    (ub[4], 5)):
        root.emplace(vals, b)
    child = O.Ibf(3, 2048, 2)
    for vals, b in ((ub[1], 0), (ub[2][:15], 1), (ub[2][15:], 2)):
        child.emplace(vals, b)
    M, D = O.BIN_MERGED, O.BIN_DELETED
    tables = dict(geoms=[(6, 2048, 2), (3, 2048, 2)], word_offsets=[0, root.words.size], tb_offsets=[0, 6, 9],
                  next_ibf=[0, 0, 1, 0, 0, 0, 1, 1, 1], user_bin=[0, 0, M, 3, D, 4, 1, 2, 2], number_of_user_bins=5,
                  arena=np.concatenate([root.words, child.words]))
    return tables, ub


def _device_hibf(t):
    from hibf_amd.device import DeviceHibf
    return DeviceHibf(t["geoms"], t["word_offsets"], t["tb_offsets"], t["next_ibf"], t["user_bin"],
                      t["number_of_user_bins"], arena=t["arena"])


def _large_ibf_tree(root_bins: int, seed: int):
    """A hand-built HIBF past the LDS budget of the descent: a root IBF of
    root_bins technical bins (user bins of one TB, a few split over two, three
    merged bins at the end), children of 700, 100 and 64 bins, and a
    grandchild of 300 bins below the first child.  Each user bin holds 12
    random values; a merged bin holds every value below it."""
    rng = np.random.default_rng(seed)
    M = O.BIN_MERGED
    root_ub = []
    u = 0
    while len(root_ub) < root_bins - 3:
        split = rng.random() < 0.05 and len(root_ub) < root_bins - 4
        root_ub += [u, u] if split else [u]
        u += 1
    layout = [root_ub + [M, M, M]]
    children = {}
    for size in (700, 100, 64, 300):
        n = size - 1 if size == 700 else size  # child 1's last bin is merged: the grandchild
        layout.append(list(range(u, u + n)) + ([M] if n < size else []))
        u += n
    children = {(0, root_bins - 3): 1, (0, root_bins - 2): 2, (0, root_bins - 1): 3, (1, 699): 4}
    n_ub = u
    vals = [rng.integers(0, 2**63, size=12, dtype=np.int64).astype(np.uint64) for _ in range(n_ub)]

    def below(i):
        out = []
        for b, x in enumerate(layout[i]):
            out += [vals[x]] if x != M else [below(children[(i, b)])]
        return np.concatenate(out)

    geoms, words, offs, tb_offs, nxt, ubs = [], [], [], [0], [], []
    at = 0
    for i, row in enumerate(layout):
        g = O.Ibf(len(row), 4096, 2)
        for b, x in enumerate(row):
            g.emplace(vals[x] if x != M else below(children[(i, b)]), b)
        geoms.append((len(row), 4096, 2))
        offs.append(at)
        w = np.concatenate([g.words, np.zeros((-g.words.size) % 4, dtype=np.uint64)])
        words.append(w)
        at += w.size
        tb_offs.append(tb_offs[-1] + len(row))
        nxt += [children.get((i, b), i) for b in range(len(row))]
        ubs += row
    t = dict(geoms=geoms, word_offsets=offs, tb_offsets=tb_offs, next_ibf=nxt, user_bin=ubs,
             number_of_user_bins=n_ub, arena=np.concatenate(words))
    return t, vals


@pytest.mark.parametrize("root_bins,children", [(9000, True), (12345, False)])
def test_large_ibfs_descend_with_state_in_device_memory(root_bins, children, descent_kernel):
    """IBFs of more than 8,192 technical bins (rows of 141 and 193 words,
    counted 128 words at a time) and trees whose counters and event lists
    exceed the LDS budget: membership lists in depth-first order and counts
    at every width against the oracle, t = 0 included (every user bin)."""
    from hibf_amd.device import to_numpy
    _say(f"large IBFs root_bins={root_bins} children={children} kernel={descent_kernel}")
    t, vals = _large_ibf_tree(root_bins, root_bins)
    if not children:  # the root alone: its merged bins become deleted bins
        n = t["geoms"][0][0]
        t = dict(geoms=t["geoms"][:1], word_offsets=[0], tb_offsets=[0, n],
                 next_ibf=[0] * n, user_bin=[x if x != O.BIN_MERGED else O.BIN_DELETED for x in t["user_bin"][:n]],
                 number_of_user_bins=max(x for x in t["user_bin"][:n] if x < O.BIN_DELETED) + 1,
                 arena=t["arena"][:t["word_offsets"][1] if len(t["word_offsets"]) > 1 else None])
    dev = _device_hibf(t)
    ref = oracle_of(dev)
    rng = np.random.default_rng(5)
    n_ub = t["number_of_user_bins"]
    qs, ts = [], []
    for k in range(40):
        u = int(rng.integers(0, n_ub)) if k % 4 else n_ub - 1 - int(rng.integers(0, 400))  # deep user bins too
        u = min(u, len(vals) - 1)
        qs.append(np.concatenate([vals[u], rng.integers(0, 2**63, size=4, dtype=np.int64).astype(np.uint64)]))
        ts.append(10)
    qs += [vals[3], np.zeros(0, dtype=np.uint64)]
    ts += [0, 0]
    got = _membership_lists(dev, qs, ts)
    for q, th, g in zip(qs, ts, got):
        assert g == ref.membership_for(q, th)
    assert sorted(set(got[-2])) == list(range(n_ub))  # t = 0: every user bin
    values, offsets = _csr(qs)
    cthr = [max(1, th) for th in ts]
    for width in (1, 2, 4, 8):
        c = to_numpy(dev.bulk_count(values, offsets, cthr, counter_bytes=width))
        assert np.array_equal(c, ref.bulk_count_batch(values, offsets, cthr, width)), width


def test_deleted_bin_hits_carry_into_the_next_run(descent_kernel):
    """A deleted TB adds its hits to the running sum without ending a run
    (hierarchical_interleaved_bloom_filter.hpp:331-340, 475-484): on the tiny
    tree, hits in deleted TB 4 count towards UB4 (TB 5)."""
    from hibf_amd.device import to_numpy
    _say(f"deleted carry kernel={descent_kernel}")
    t, ub = _tiny_tree_with_deleted_bin()
    dev = _device_hibf(t)
    q = np.concatenate([ub[5][:4], ub[4][:2]])
    assert _membership_lists(dev, [q, ub[1][:8], np.concatenate([ub[4][:6], ub[0][:6], ub[3][:6]])], [6, 5, 6]) == \
        [[4], [1], [0, 3, 4]]
    values, offsets = _csr([q])
    assert to_numpy(dev.bulk_count(values, offsets, [6])).tolist() == [[0, 0, 0, 0, 6]]
    # TB 1 deleted as well: UB0's run ends at TB 0, and TB 1's hits go to the merged bin's sum.
    t["user_bin"] = list(t["user_bin"])
    t["user_bin"][1] = O.BIN_DELETED
    dev = _device_hibf(t)
    ref = oracle_of(dev)
    qs = [np.concatenate([ub[0][20:26], ub[1][:3]]), ub[0][:10], np.concatenate([ub[0][20:40], ub[2][:2]]),
          np.concatenate([ub[5][:4], ub[4][:2], ub[0][25:30]])]
    assert _membership_lists(dev, qs, [8, 5, 6, 6]) == [ref.membership_for(q, th) for q, th in zip(qs, [8, 5, 6, 6])]
    # The carry is seen in the visits: 6 hits in TB 1 + 3 in TB 2 reach t = 8 and the child IBF is visited
    # (UB1's 3 hits stay below t); with one UB1 value the sum is 7 and only the root is visited.
    from hibf_amd.device import Agent
    for extra, visits in ((3, 2), (1, 1)):
        agent = Agent()
        agent.stats(reset=True)
        v1, o1 = _csr([np.concatenate([ub[0][20:26], ub[1][:extra]])])
        ids, _ = dev.membership_for(v1, o1, [8], agent)
        assert ids.numel() == 0
        assert agent.stats()["visits"] == visits, extra
    values, offsets = _csr(qs)
    assert np.array_equal(to_numpy(dev.bulk_count(values, offsets, [3, 3, 3, 3])),
                          ref.bulk_count_batch(values, offsets, [3, 3, 3, 3], 2))


def test_runs_cut_by_deleted_bins(descent_kernel):
    """One IBF [UB0, deleted, UB0, deleted, UB0, UB1]: threshold 0 reports UB0
    once per run (more ids than user bins), and bulk_count keeps the last run
    sum that reached the threshold (hierarchical_interleaved_bloom_filter.hpp:488-491),
    as the oracle's brute-forced answers say."""
    from hibf_amd.device import to_numpy
    _say(f"runs cut by deleted bins kernel={descent_kernel}")
    rng = np.random.default_rng(5)
    parts = [rng.integers(0, 2**63, size=20, dtype=np.int64).astype(np.uint64) for _ in range(6)]
    ibf = O.Ibf(6, 4096, 2)
    for b, p in enumerate(parts):
        ibf.emplace(p, b)
    D = O.BIN_DELETED
    dev = _device_hibf(dict(geoms=[(6, 4096, 2)], word_offsets=[0], tb_offsets=[0, 6], next_ibf=[0] * 6,
                            user_bin=[0, D, 0, D, 0, 1], number_of_user_bins=2, arena=ibf.words))
    q = np.concatenate([parts[0][:7], parts[1][:2], parts[2][:3], parts[4][:4], parts[5][:1]])
    assert _membership_lists(dev, [parts[5][:1], q], [0, 5]) == [[0, 0, 0, 1], [0, 0]]
    values, offsets = _csr([q, q, q])
    assert to_numpy(dev.bulk_count(values, offsets, [5, 4, 8])).tolist() == [[5, 0], [4, 0], [0, 0]]


def test_bulk_count_refuses_a_user_bin_in_two_ibfs():
    """Tables that list one user bin in two IBFs (never built by the
    reference) answer membership_for, and bulk_count fails loudly."""
    from hibf_amd import native
    _say("user bin in two IBFs")
    t, ub = _tiny_tree_with_deleted_bin()
    t["user_bin"] = list(t["user_bin"])
    t["user_bin"][5] = 1   # UB1 also lives in the child IBF
    dev = _device_hibf(t)
    ref = oracle_of(dev)
    q = np.concatenate([ub[4][:6], ub[1][:6]])
    assert _membership_lists(dev, [q], [5]) == [ref.membership_for(q, 5)]
    values, offsets = _csr([q])
    with pytest.raises(native.HibfGpuError) as e:
        dev.bulk_count(values, offsets, [5])
    assert e.value.code == native.ERR_UNSUPPORTED


def test_hibf_with_deleted_bins(descent_kernel):
    """A built HIBF with about 15 % of its user-bin TBs marked deleted (their
    bits kept), including ends and middles of split runs and TBs just before
    merged bins: membership and counts equal the oracle's on the same tables."""
    from hibf_amd.device import to_numpy
    _say(f"deleted bins kernel={descent_kernel}")
    rng = np.random.default_rng(41)
    ubs = _random_user_bins(rng, 400, 20, 400)
    built = _build_cached_deleted(ubs)
    t = dict(built)
    user_bin = t["user_bin"].copy()
    regular = np.nonzero(user_bin < np.uint64(O.BIN_DELETED))[0]
    user_bin[rng.choice(regular, size=len(regular) * 15 // 100, replace=False)] = np.uint64(O.BIN_DELETED)
    starts = set(int(x) for x in t["tb_offsets"][:-1])
    before_merged = [b - 1 for b in range(1, len(user_bin)) if user_bin[b] == np.uint64(O.BIN_MERGED)
                     and b not in starts and user_bin[b - 1] < np.uint64(O.BIN_DELETED)]
    user_bin[before_merged[:3]] = np.uint64(O.BIN_DELETED)
    t["user_bin"] = user_bin
    dev = _device_hibf(t)
    ref = oracle_of(dev)
    qs = _mixed_queries(rng, ubs, 80, 150)
    thresholds = [int(x) for x in rng.integers(1, 70, size=len(qs))]
    thresholds[0] = 0
    got = _membership_lists(dev, qs, thresholds)
    for q, th, g in zip(qs, thresholds, got):
        assert g == ref.membership_for(q, th)
    values, offsets = _csr(qs)
    cthr = [max(1, th) for th in thresholds]
    for width in (1, 2):
        c = to_numpy(dev.bulk_count(values, offsets, cthr, counter_bytes=width))
        assert np.array_equal(c, ref.bulk_count_batch(values, offsets, cthr, width)), width


_DELETED_CACHE = {}


def _build_cached_deleted(ubs):
    """Tables and arena of the 400-UB HIBF (built once for all descent kernels)."""
    if "t" not in _DELETED_CACHE:
        from hibf_amd.device import DeviceHibf
        dev = DeviceHibf.build(ubs, tmax=64)
        tab = dev.tables()
        _DELETED_CACHE["t"] = dict(geoms=tab["geoms"], word_offsets=tab["word_offsets"],
                                   tb_offsets=tab["tb_offsets"], next_ibf=tab["next_ibf"],
                                   user_bin=tab["user_bin"], number_of_user_bins=tab["n_user_bins"],
                                   arena=dev.download_arena())
    return _DELETED_CACHE["t"]


def _same_hibf(a, b):
    ta, tb = a.tables(), b.tables()
    assert ta["geoms"] == tb["geoms"]
    for k in ("word_offsets", "tb_offsets", "next_ibf", "user_bin"):
        assert np.array_equal(ta[k], tb[k]), k
    assert np.array_equal(a.download_arena(), b.download_arena())


def test_build_layout_sees_hyperloglog_estimates():
    """The builders hand the layout the user bins' HyperLogLog estimates, as
    the reference's build does (src/hierarchical_interleaved_bloom_filter.cpp:
    215-230): with disable_estimate_union each build equals the build from the
    layout of the oracle's estimates (12-bit sketches); by default (union
    estimates and rearrangement) it equals the build from the layout computed
    with the user bins' sketches, for explicit and for synthetic user bins."""
    from hibf_amd import native
    from hibf_amd.device import DeviceHibf, hll_sketches, hll_sketches_synthetic
    from hibf_amd.synth import synth_values
    _say("layout from HLL estimates")
    rng = np.random.default_rng(17)
    cfg = dict(tmax=64)
    ubs = _random_user_bins(rng, 300, 200, 20000)
    counts = [int(O.hll_estimate(O.hll_sketch(np.unique(u), 12))) for u in ubs]
    assert any(c != np.unique(u).size for c, u in zip(counts, ubs))   # the estimates are not the exact counts
    sums = dict(cfg, disable_estimate_union=True)
    text = native.compute_layout(native.make_config(number_of_user_bins=len(ubs), **sums), counts)
    _same_hibf(DeviceHibf.build(ubs, **sums), DeviceHibf.build(ubs, layout_text=text, **sums))
    c = native.validate_config(native.make_config(number_of_user_bins=len(ubs), **cfg))
    text = native.compute_layout_sketched(c, counts, hll_sketches([np.unique(u) for u in ubs]).data_ptr())
    assert text != native.compute_layout(c, counts)
    _same_hibf(DeviceHibf.build(ubs, **cfg), DeviceHibf.build(ubs, layout_text=text, **cfg))
    sizes = rng.integers(100, 30000, size=500)
    seed = 5
    counts = [int(O.hll_estimate(O.hll_sketch(synth_values(seed, u, np.arange(n, dtype=np.uint64)), 12)))
              for u, n in enumerate(sizes)]
    text = native.compute_layout(native.make_config(number_of_user_bins=len(sizes), **sums), counts)
    _same_hibf(DeviceHibf.build_synthetic(sizes, seed, **sums),
               DeviceHibf.build_synthetic(sizes, seed, layout_text=text, **sums))
    c = native.validate_config(native.make_config(number_of_user_bins=len(sizes), **cfg))
    text = native.compute_layout_sketched(c, counts, hll_sketches_synthetic(sizes, seed).data_ptr())
    _same_hibf(DeviceHibf.build_synthetic(sizes, seed, **cfg),
               DeviceHibf.build_synthetic(sizes, seed, layout_text=text, **cfg))


def test_hibf_synthetic_build_matches_oracle():
    from hibf_amd.device import DeviceHibf, to_numpy
    from hibf_amd.synth import synth_values
    _say("synthetic build")
    rng = np.random.default_rng(3)
    sizes = np.clip(rng.lognormal(np.log(300), 1.0, size=700).astype(np.int64), 20, 5000)
    seed = 77
    dev = DeviceHibf.build_synthetic(sizes, seed, tmax=64)
    ubs = [synth_values(seed, u, np.arange(s, dtype=np.uint64)) for u, s in enumerate(sizes)]
    # split bins are chunked in the reference's set order of the generated values
    chunks = synthetic_split_chunks(dev, seed, sizes)
    assert len(chunks) > 0
    assert np.array_equal(dev.download_arena(), expected_arena(dev, ubs, keep_order=True, split_chunks=chunks))
    ref = oracle_of(dev)
    qs = _mixed_queries(rng, ubs, 80, 200)
    ts = [int(0.7 * len(q)) // 2 for q in qs]
    assert _membership_lists(dev, qs, ts) == [ref.membership_for(q, t) for q, t in zip(qs, ts)]
    values, offsets = _csr(qs)
    assert np.array_equal(to_numpy(dev.bulk_count(values, offsets, ts)), ref.bulk_count_batch(values, offsets, ts))


@pytest.mark.parametrize("pinned", [False, True], ids=["pageable", "pinned"])
def test_serving_results_equal_alone_and_in_the_mix(pinned):
    """Each request's ids from the serving loop (batched with others, several
    streams in flight) equal its ids when it is answered alone; query values
    handed over in pageable memory (staged) or in a pinned buffer (direct)."""
    import torch
    from hibf_amd.device import DeviceHibf
    _say(f"serving mixed vs alone, {'pinned' if pinned else 'pageable'} values")
    rng = np.random.default_rng(21)
    ubs = _random_user_bins(rng, 400, 20, 300)
    dev = DeviceHibf.build(ubs, tmax=64)
    qs = _mixed_queries(rng, ubs, 3000, 100)
    ts = np.array([int(0.3 * len(q)) for q in qs], dtype=np.uint16)
    ts[::97] = 1   # some requests with many hits
    values, offsets = _csr(qs)
    arrival = np.cumsum(rng.exponential(1 / 200000, size=len(qs)))
    arrival -= arrival[0]
    v_in = values
    if pinned:
        v_in = torch.empty(values.size, dtype=torch.int64, pin_memory=True)
        v_in.numpy()[:] = values.view(np.int64)
    lat, ids, cnt, st = dev.serve(v_in, offsets, ts, arrival, window_us=300, streams=3, max_batch=256,
                                  results_per_query=64)
    assert (lat > 0).all() and st["batches"] > 1
    alone = [ _membership_lists(dev, [qs[i]], [int(ts[i])])[0] for i in range(0, len(qs), 37)]
    for k, i in enumerate(range(0, len(qs), 37)):
        assert int(cnt[i]) == len(alone[k])
        assert ids[i, :min(64, int(cnt[i]))].tolist() == alone[k][:64]
    batch = _membership_lists(dev, qs, ts)
    for i in range(len(qs)):
        assert int(cnt[i]) == len(batch[i]) and ids[i, :min(64, int(cnt[i]))].tolist() == batch[i][:64]


def test_serving_large_staged_batches():
    """Batches above 8 MiB of pageable query values are staged by several host
    threads; every request's ids still equal the batched membership result."""
    from hibf_amd.device import DeviceHibf
    _say("serving large staged batches")
    rng = np.random.default_rng(22)
    ubs = _random_user_bins(rng, 300, 50, 400)
    dev = DeviceHibf.build(ubs, tmax=64)
    qs = _mixed_queries(rng, ubs, 12000, 200)
    ts = np.array([int(0.5 * len(q)) for q in qs], dtype=np.uint16)
    values, offsets = _csr(qs)
    arrival = np.cumsum(rng.exponential(1 / 4e6, size=len(qs)))
    arrival -= arrival[0]
    lat, ids, cnt, st = dev.serve(values, offsets, ts, arrival, window_us=5000, streams=2, max_batch=65536,
                                  max_batch_values=1 << 23, results_per_query=8)
    assert int(offsets[-1]) * 8 / st["batches"] > (8 << 20) / 4  # batches of megabytes
    batch = _membership_lists(dev, qs, ts)
    for i in range(len(qs)):
        assert int(cnt[i]) == len(batch[i]) and ids[i, :min(8, int(cnt[i]))].tolist() == batch[i][:8]


def _flat_layout(n_ub, split0=1):
    """Layout text with every user bin in the top-level IBF, user bin 0 split
    over `split0` technical bins (src/layout/layout.cpp:125-136 format)."""
    lines = ["#TOP_LEVEL_IBF fullest_technical_bin_idx:0",
             "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS",
             f"0\t0\t{split0}"]
    lines += [f"{u}\t{u + split0 - 1}\t1" for u in range(1, n_ub)]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("n_ub", [3000, 5000])
def test_wide_single_ibf_hibf(n_ub, descent_kernel):
    """One IBF of 3,002 / 5,002 technical bins (47 / 79 words per row): the
    widest counting shapes of the descent, with a split user bin."""
    from hibf_amd.device import DeviceHibf, to_numpy
    _say(f"wide single-IBF HIBF n_ub={n_ub} kernel={descent_kernel}")
    rng = np.random.default_rng(n_ub)
    ubs = _random_user_bins(rng, n_ub, 5, 60)
    dev = DeviceHibf.build(ubs, layout_text=_flat_layout(n_ub, 3), tmax=8192)
    t = dev.tables()
    assert len(t["geoms"]) == 1 and t["geoms"][0][0] == n_ub + 2
    ref = oracle_of(dev)
    qs = _mixed_queries(rng, ubs, 80, 120) + [np.zeros(0, dtype=np.uint64)]
    thresholds = [int(x) for x in rng.integers(1, 60, size=len(qs))]
    got = _membership_lists(dev, qs, thresholds)
    for q, th, g in zip(qs, thresholds, got):
        assert g == ref.membership_for(q, th)
    values, offsets = _csr(qs)
    counts = to_numpy(dev.bulk_count(values, offsets, thresholds, 2))
    assert np.array_equal(counts, ref.bulk_count_batch(values, offsets, thresholds, 2))


def test_hibf_long_queries_wrap(descent_kernel):
    """Queries longer than 65,535 values: the membership sums are uint16 and
    the counting sums value_t, and both wrap as the reference's do
    (hierarchical_interleaved_bloom_filter.hpp:329-354, :462-496)."""
    from hibf_amd.device import DeviceHibf, to_numpy
    _say(f"hibf long queries kernel={descent_kernel}")
    rng = np.random.default_rng(77)
    ubs = _random_user_bins(rng, 300, 20, 200)
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    qs = [np.full(70000, ubs[5][0], dtype=np.uint64),
          np.full(65536, ubs[7][1], dtype=np.uint64),
          np.full(65535, ubs[9][2], dtype=np.uint64),
          np.concatenate([ubs[3][rng.integers(0, ubs[3].size, size=66000)],
                          rng.integers(0, 2**63, size=1000, dtype=np.int64).astype(np.uint64)])]
    for th in (1, 100, 4465, 5000, 65535):
        got = _membership_lists(dev, qs, [th] * len(qs))
        for q, g in zip(qs, got):
            assert g == ref.membership_for(q, th), th
    values, offsets = _csr(qs)
    for width in (1, 2, 4, 8):
        for th in (1, 3000):
            got = to_numpy(dev.bulk_count(values, offsets, [th] * len(qs), width))
            assert np.array_equal(got, ref.bulk_count_batch(values, offsets, [th] * len(qs), width)), (width, th)


def test_split_bins_in_reference_set_order_with_duplicates_and_given_order():
    """User bins handed over unsorted and with repeated values: the reference's
    k-mer set sees them in that order (duplicates ignored), and every split
    user bin's technical bins hold the chunks of that set's iteration order.
    With build.split_order = 0 the same user bins are chunked by sorted value
    ranges instead."""
    from hibf_amd import native
    from hibf_amd.device import DeviceHibf
    _say("split bins: reference set order, duplicates, given order")
    rng = np.random.default_rng(88)
    ubs = []
    for n in rng.integers(30, 3000, size=300):
        v = rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
        ubs.append(np.concatenate([v, v[rng.integers(0, v.size, size=int(n) // 3)]]))  # repeats, unsorted
    ubs[0] = np.concatenate([ubs[0], rng.integers(0, 2**63, size=40000, dtype=np.int64).astype(np.uint64)])
    dev = DeviceHibf.build(ubs, tmax=64)
    chunks = explicit_split_chunks(dev, ubs)
    assert len(chunks) > 0
    arena = dev.download_arena()
    assert np.array_equal(arena, expected_arena(dev, ubs, split_chunks=chunks))
    try:
        native.tune_set("build.split_order", 0)
        dev0 = DeviceHibf.build(ubs, layout_text=dev.layout_text(), tmax=64)
    finally:
        native.tune_set("build.split_order", 1)
    arena0 = dev0.download_arena()
    assert np.array_equal(arena0, expected_arena(dev0, ubs))
    assert not np.array_equal(arena0, arena)
    ref = oracle_of(dev)
    qs = _mixed_queries(rng, ubs, 50, 150)
    assert _membership_lists(dev, qs, [60] * len(qs)) == [ref.membership_for(q, 60) for q in qs]
