"""GPU parity of the flat IBF path (emplace, clear, bulk_contains, bulk_count,
membership_for) against the oracle, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GEOMS = [  # (bins, bin_size, hash_funs)
    (1, 64, 1), (7, 1000, 2), (63, 513, 2), (64, 4096, 3), (65, 777, 5), (127, 1024, 2), (130, 2000, 2),
    (1000, 300, 4), (4096, 3000, 2), (8192, 2048, 3), (8190, 700, 3), (9000, 500, 2), (16384, 300, 2),
    (20000, 129, 1),
]


def _say(msg):
    print(msg, flush=True)


def _rand_u64(rng, n):
    return rng.integers(0, 2**63, size=n, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, size=n).astype(np.uint64)


def _pair(bins, bin_size, h, n, seed):
    from hibf_amd.device import DeviceIbf
    rng = np.random.default_rng(seed)
    vals = _rand_u64(rng, n)
    bidx = rng.integers(0, bins, size=n).astype(np.uint64)
    ref = O.Ibf(bins, bin_size, h)
    ref.emplace(vals, bidx)
    dev = DeviceIbf(bins, bin_size, h)
    dev.emplace(vals, bidx)
    return ref, dev, vals, bidx, rng


def _queries(rng, vals, lengths, frac_in=0.7):
    qs = []
    for L in lengths:
        k = int(L * frac_in)
        q = np.concatenate([vals[rng.integers(0, len(vals), size=k)], _rand_u64(rng, L - k)])
        qs.append(q)
    return qs


@pytest.mark.parametrize("bins,bin_size,h", GEOMS)
def test_emplace_matrix_identical(bins, bin_size, h):
    _say(f"emplace bins={bins} bin_size={bin_size} h={h}")
    ref, dev, *_ = _pair(bins, bin_size, h, 20000, bins)
    assert np.array_equal(dev.download(), ref.words)


def test_emplace_single_bin_and_geometry():
    from hibf_amd.device import DeviceIbf
    _say("emplace constant bin")
    ref = O.Ibf(73, 1019, 3)
    dev = DeviceIbf(73, 1019, 3)
    assert (dev.bin_count(), dev.bin_size(), dev.hash_function_count(), dev.bit_size()) == (73, 1019, 3, 130_432)
    vals = np.arange(5000, dtype=np.uint64)
    ref.emplace(vals, 72)
    dev.emplace(vals, 72)
    assert np.array_equal(dev.download(), ref.words)


@pytest.mark.parametrize("bins,bin_size,h", GEOMS)
def test_bulk_contains(bins, bin_size, h):
    from hibf_amd.device import to_numpy
    _say(f"contains bins={bins} bin_size={bin_size} h={h}")
    ref, dev, vals, _, rng = _pair(bins, bin_size, h, 5000, bins + 1)
    q = np.concatenate([vals[:501], _rand_u64(rng, 500)])  # 1001 values: a ragged last wave
    got = to_numpy(dev.bulk_contains(q))
    want = np.stack([ref.bulk_contains(int(v)) for v in q])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bins,bin_size,h", GEOMS)
@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_bulk_count(bins, bin_size, h, width):
    from hibf_amd.device import csr, to_numpy
    _say(f"count bins={bins} bin_size={bin_size} h={h} width={width}")
    ref, dev, vals, _, rng = _pair(bins, bin_size, h, 30000, bins + 2)
    lengths = [0, 1, 2, 63, 64, 65, 127, 250, 300, 1, 0, 700]
    qs = _queries(rng, vals, lengths)
    values, offsets = csr(qs)
    got = to_numpy(dev.bulk_count(values, offsets, width))
    want = ref.bulk_count_batch(values, offsets, width)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_bulk_count_wraps_like_counting_vector():
    from hibf_amd.device import DeviceIbf, csr, to_numpy
    _say("count wrap uint8/uint16")
    dev = DeviceIbf(3, 4096, 2)
    dev.emplace([42], 1)
    values, offsets = csr([np.full(300, 42, dtype=np.uint64), np.full(70000, 42, dtype=np.uint64)])
    assert to_numpy(dev.bulk_count(values, offsets, 1)).tolist() == [[0, 300 % 256, 0], [0, 70000 % 256, 0]]
    assert to_numpy(dev.bulk_count(values, offsets, 2)).tolist() == [[0, 300, 0], [0, 70000 % 65536, 0]]
    assert to_numpy(dev.bulk_count(values, offsets, 8)).tolist() == [[0, 300, 0], [0, 70000, 0]]


@pytest.mark.parametrize("bins", [8192, 16384, 8190])
def test_bulk_count_long_queries_wide_rows(bins):
    """Queries longer than 65,535 values on wide rows: counters are flushed per
    segment and must still wrap exactly like counting_vector<T>."""
    from hibf_amd.device import DeviceIbf, csr, to_numpy
    _say(f"count long queries bins={bins}")
    ref = O.Ibf(bins, 2048, 2)
    dev = DeviceIbf(bins, 2048, 2)
    for b in (0, 5, bins - 1):
        ref.emplace([42, 43], b)
        dev.emplace([42, 43], b)
    rng = np.random.default_rng(bins)
    qs = [np.full(70000, 42, dtype=np.uint64),
          np.concatenate([np.full(140000, 43, dtype=np.uint64), _rand_u64(rng, 1000)]),
          np.full(65535, 42, dtype=np.uint64), np.full(65536, 43, dtype=np.uint64)]
    values, offsets = csr(qs)
    for width in (1, 2, 4, 8):
        got = to_numpy(dev.bulk_count(values, offsets, width))
        assert np.array_equal(got, ref.bulk_count_batch(values, offsets, width)), width


@pytest.fixture
def knobs():
    """Sets launch-shape knobs for one test and restores the defaults after it."""
    from hibf_amd import native
    keys = ("count.wide_unroll", "count.wide_waves", "count.counter_bits", "count.nt_stores", "contains.unroll")
    saved = {k: native.tune_get(k) for k in keys}

    def set_(**kv):
        for k, v in kv.items():
            native.tune_set(k.replace("__", "."), v)
    yield set_
    for k, v in saved.items():
        native.tune_set(k, v)


@pytest.mark.parametrize("nt_stores", [0, 1])
def test_counter_store_forms(nt_stores, knobs):
    """Plain and non-temporal counter stores give the oracle's counts (flat
    wide kernel, segmented long query included) and the same HIBF counts."""
    from hibf_amd.device import DeviceHibf, csr, to_numpy
    from hibf_helpers import oracle_of
    _say(f"counter stores nt={nt_stores}")
    knobs(count__nt_stores=nt_stores)
    ref, dev, vals, _, rng = _pair(8192, 1024, 3, 20000, 5)
    qs = _queries(rng, vals, [0, 1, 250, 300, 1000]) + [np.full(70000, vals[7], dtype=np.uint64)]
    values, offsets = csr(qs)
    for width in (1, 2, 4):
        assert np.array_equal(to_numpy(dev.bulk_count(values, offsets, width)),
                              ref.bulk_count_batch(values, offsets, width)), width
    ubs = [_rand_u64(rng, int(n)) for n in rng.integers(20, 200, size=301)]   # odd row length in bytes
    hb = DeviceHibf.build(ubs, tmax=64)
    href = oracle_of(hb)
    hq = [np.concatenate([ubs[int(i)][:40], _rand_u64(rng, 10)]) for i in rng.integers(0, 301, size=50)]
    hv, ho = csr(hq)
    for width in (1, 2, 8):
        assert np.array_equal(to_numpy(hb.bulk_count(hv, ho, [20] * 50, counter_bytes=width)),
                              href.bulk_count_batch(hv, ho, [20] * 50, width)), width


@pytest.mark.parametrize("unroll", [1, 2, 4, 8])
@pytest.mark.parametrize("nt_stores", [0, 1])
@pytest.mark.parametrize("bins", [8192, 9000])
def test_contains_store_forms(unroll, nt_stores, bins, knobs):
    """bulk_contains rows are the oracle's for 1, 2, 4 and 8 values in flight
    per wave, with plain and non-temporal stores, into a 16-byte aligned output
    and into one 8 bytes off (the 8-byte path), for an even (128) and an odd
    (141) row width in words; 633 values leave a ragged last group."""
    import torch
    _say(f"contains stores unroll={unroll} nt={nt_stores} bins={bins}")
    knobs(count__nt_stores=nt_stores, contains__unroll=unroll)
    ref, dev, vals, _, rng = _pair(bins, 1024, 3, 20000, 11)
    q = np.concatenate([vals[:333], _rand_u64(rng, 300)])
    want = np.stack([ref.bulk_contains(int(v)) for v in q])
    bw = want.shape[1]
    qd = torch.from_numpy(q.view(np.int64)).cuda()
    for shift in (0, 1):
        buf = torch.full((q.size * bw + 1,), -1, dtype=torch.int64, device="cuda")
        out = buf[shift:shift + q.size * bw]
        dev.bulk_contains_into(qd, out)
        got = out.cpu().numpy().view(np.uint64).reshape(q.size, bw)
        assert np.array_equal(got, want), shift
        assert int(buf[q.size * bw if shift == 0 else 0]) == -1, shift  # nothing past the rows


@pytest.mark.parametrize("h", [2, 3, 4])
@pytest.mark.parametrize("counter_bits", [8, 16])
@pytest.mark.parametrize("unroll,waves", [(1, 1), (2, 4), (4, 2), (8, 4), (2, 1), (4, 4)])
def test_bulk_count_tile_shapes(h, counter_bits, unroll, waves, knobs):
    """Every tile shape of the wide count kernel (values in flight, waves per
    block, 8/16-bit LDS counters, compile-time and run-time hash counts) gives
    the oracle's counts; lengths straddle the 255- and 65,535-value segments."""
    from hibf_amd.device import csr, to_numpy
    _say(f"count tile h={h} bits={counter_bits} unroll={unroll} waves={waves}")
    knobs(count__counter_bits=counter_bits, count__wide_unroll=unroll, count__wide_waves=waves)
    ref, dev, vals, _, rng = _pair(8192, 1024, h, 20000, 77 + h)
    qs = _queries(rng, vals, [0, 1, 63, 254, 255, 256, 300, 511, 1200])
    qs += [np.full(300, vals[3], dtype=np.uint64), np.full(66000, vals[5], dtype=np.uint64)]
    values, offsets = csr(qs)
    for width in (1, 2, 8):
        got = to_numpy(dev.bulk_count(values, offsets, width))
        assert np.array_equal(got, ref.bulk_count_batch(values, offsets, width)), width


@pytest.mark.parametrize("bins,bin_size,h", [(12, 8192, 2), (200, 3000, 3), (8192, 1024, 2), (9000, 300, 2)])
def test_membership_for(bins, bin_size, h):
    from hibf_amd.device import csr, to_numpy
    _say(f"membership bins={bins}")
    ref, dev, vals, _, rng = _pair(bins, bin_size, h, 40000, bins + 3)
    qs = _queries(rng, vals, [0, 5, 40, 100, 250, 33])
    values, offsets = csr(qs)
    thresholds = [0, 1, 2, 3, 4, 1]
    ids, hoff = dev.membership_for(values, offsets, thresholds)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    for i, (q, t) in enumerate(zip(qs, thresholds)):
        assert ids[hoff[i]:hoff[i + 1]].tolist() == ref.membership_for(q, t)


@pytest.mark.parametrize("bins,bin_size,h", [(8192, 1024, 3), (1003, 700, 2), (64, 2000, 2)])
def test_membership_for_chunked(bins, bin_size, h):
    """An agent limited to 7 queries per launch splits 60 queries into 9 chunks
    whose offsets continue on the device; hits (threshold 0 = every bin,
    empty queries, thresholds above the length) equal the oracle's and the
    one-chunk agent's."""
    from hibf_amd.device import Agent, csr, to_numpy
    _say(f"membership chunked bins={bins}")
    ref, dev, vals, _, rng = _pair(bins, bin_size, h, 30000, bins + 11)
    qs = _queries(rng, vals, [int(n) for n in rng.integers(0, 120, size=60)])
    qs[4] = np.zeros(0, dtype=np.uint64)
    values, offsets = csr(qs)
    thresholds = rng.integers(0, 60, size=len(qs)).astype(np.uint16)
    thresholds[[0, 9, 30]] = 0
    thresholds[17] = 500
    ids, hoff = dev.membership_for(values, offsets, thresholds, Agent(7))
    ids1, hoff1 = dev.membership_for(values, offsets, thresholds, Agent())
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    assert np.array_equal(ids, to_numpy(ids1)) and np.array_equal(hoff, to_numpy(hoff1))
    assert hoff[0] == 0 and hoff[-1] == ids.size
    for i, (q, t) in enumerate(zip(qs, thresholds)):
        assert ids[hoff[i]:hoff[i + 1]].tolist() == ref.membership_for(q, int(t)), i
    assert hoff[1] - hoff[0] == bins  # threshold 0 on query 0


def test_clear_matches_oracle():
    from hibf_amd.device import csr, to_numpy
    _say("clear")
    ref, dev, vals, _, rng = _pair(150, 2000, 2, 20000, 99)
    for bins in ([17], [0, 63, 64, 149]):
        ref_words = ref.words
        for b in bins:
            ref.clear(b)
        dev.clear(bins)
        assert np.array_equal(dev.download(), ref.words)
    values, offsets = csr(_queries(rng, vals, [100, 200]))
    assert np.array_equal(to_numpy(dev.bulk_count(values, offsets)), ref.bulk_count_batch(values, offsets))


@pytest.mark.parametrize("bins,bin_size,h", [(64, 97, 2), (130, 61, 3), (5, 4096, 5), (300, 1, 2), (70, 2, 1)])
def test_emplace_tracked_occupancy_matches_oracle(bins, bin_size, h):
    """emplace with track_occupancy on the device: the occupancy equals the
    oracle's sequential emplace (src/interleaved_bloom_filter.cpp:119-140) over
    pairs with repeated values, colliding rows (small bin sizes), out-of-range
    bins, a whole-range call into one bin and a second call on the same filter;
    the bits equal the untracked emplace's."""
    from hibf_amd.device import DeviceIbf
    _say(f"tracked emplace bins={bins} bin_size={bin_size} h={h}")
    rng = np.random.default_rng(bins * 7 + h)
    dev = DeviceIbf(bins, bin_size, h)
    ref = O.Ibf(bins, bin_size, h)
    tb = int(dev.geometry.technical_bins)
    dev.set_occupancy(np.arange(tb, dtype=np.uint64) % 3, True)  # counters continue from what the caller set
    occ = np.arange(tb, dtype=np.uint64) % 3
    for call in range(3):
        vals = rng.integers(0, 2**63, size=4000, dtype=np.int64).astype(np.uint64)
        vals[1000:1500] = vals[:500]
        if call == 2:
            dev.emplace(vals[:700], bins - 1)
            ref.emplace_tracked(vals[:700], bins - 1, occ)
        else:
            bs = rng.integers(0, bins + 3, size=vals.size).astype(np.uint64)  # some bins out of range: skipped
            dev.emplace(vals, bs)
            ref.emplace_tracked(vals, bs, occ)
        got, track = dev.occupancy()
        assert track and np.array_equal(got, occ), call
        assert np.array_equal(dev.download(), ref.words), call
    dev.set_occupancy(np.zeros(tb, dtype=np.uint64), False)
    dev.emplace(vals, 0)  # untracked: counters stay
    assert not dev.occupancy()[0].any()


def test_snippet_goldens_on_gpu(goldens):
    from hibf_amd.device import DeviceIbf, csr, to_numpy
    _say("snippet goldens")

    def vals(spec):
        return np.arange(*spec["range"], dtype=np.uint64) if "range" in spec else np.asarray(spec["values"],
                                                                                             dtype=np.uint64)

    def build(case):
        d = DeviceIbf(case["bins"], case["bin_size"], case["hash_funs"])
        for e in case["emplace"]:
            d.emplace(vals(e), e["bin"])
        return d

    case = goldens["ibf_counting_agent"]
    d = build(case)
    for q in case["queries"]:
        v, o = csr([vals(q)])
        assert to_numpy(d.bulk_count(v, o, q["width"]))[0].tolist() == q["expected"]
    case = goldens["ibf_clear"]
    d = build(case)
    for step in case["steps"]:
        if step["clear"]:
            d.clear(step["clear"])
        v, o = csr([vals(step)])
        assert to_numpy(d.bulk_count(v, o))[0].tolist() == step["expected"]
    case = goldens["ibf_containment"]
    w = to_numpy(build(case).bulk_contains([case["value"]]))[0]
    assert [int((int(w[b // 64]) >> (b % 64)) & 1) for b in range(case["bins"])] == case["expected"]
    case = goldens["ibf_membership"]
    v, o = csr([case["query"]])
    ids, hoff = build(case).membership_for(v, o, [case["threshold"]])
    assert to_numpy(ids).tolist() == case["expected"]
    case = goldens["insert_iterator_ibf"]
    v, o = csr([np.arange(*case["query_range"], dtype=np.uint64)])
    assert to_numpy(build(case).bulk_count(v, o, 1))[0].tolist() == case["expected"]


@pytest.mark.parametrize("bins,bin_size,h,new_bins", [(70, 1000, 2, 130), (64, 333, 3, 65), (3, 500, 2, 64),
                                                      (130, 777, 2, 1000), (8192, 64, 3, 8193), (4000, 2000, 2, 9000)])
def test_increase_bin_number_on_device(bins, bin_size, h, new_bins):
    """Device re-stride equals the oracle's in-place re-stride and a fresh
    build with the new bin count; counts afterwards equal the oracle's."""
    from hibf_amd.device import DeviceIbf, csr, to_numpy
    _say(f"increase bins {bins} -> {new_bins}")
    ref, dev, vals, bidx, rng = _pair(bins, bin_size, h, 5000, bins + new_bins)
    same_words = (new_bins + 63) // 64 == (bins + 63) // 64
    assert dev.try_increase_bin_number_to(new_bins) == same_words
    dev.increase_bin_number_to(new_bins)
    ref.increase_bin_number_to(new_bins)
    fresh = DeviceIbf(new_bins, bin_size, h)
    fresh.emplace(vals, bidx)
    assert dev.bin_count() == new_bins and dev.geometry.bin_words == (new_bins + 63) // 64
    assert np.array_equal(dev.download(), ref.words)
    assert np.array_equal(dev.download(), fresh.download())
    qs = _queries(rng, vals, [0, 1, 100, 250])
    v, o = csr(qs)
    assert np.array_equal(to_numpy(dev.bulk_count(v, o, 2)), ref.bulk_count_batch(v, o, 2))
    with pytest.raises(Exception, match="must be >= the current number of bins"):
        dev.increase_bin_number_to(new_bins - 1)
