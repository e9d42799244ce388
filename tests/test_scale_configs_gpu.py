"""GPU parity at the sizes the path is measured on.

- Flat IBF: 8,192 technical bins x 2^23 rows (8 GiB), 3 hash functions, 2^17
  generated values per bin by the GPU bulk emplace; the whole bit matrix word
  for word against the oracle's emplace of the same values; uint16 bulk_count
  of 1,000 queries (70 % of the values from one bin) and bulk_contains.
- HIBF at scale: 32,768 user bins of log-normal sizes (median 1e5, sigma 1.5,
  clipped to [1e3, 5e7]), tmax 256, h = 2, FPR 0.05 / 0.3, on the
  reference-default layout, built on the device: the columns of the top-level
  IBF and of 24 sampled lower IBFs against the oracle's emplace; membership_for
  of 600 queries (t = uint16(0.7 L)) and bulk_count of 120 against the oracle.
- Batched form: one hipGraph batch of 65,536 log-normal-length queries,
  replayed twice with eager work between, against one launch per query and
  the oracle.
- Serving: 0.2 s of Poisson arrivals at 500,000 queries/s: each request's
  ids in the mix equal the batched answer, the answer alone and the oracle's.
"""
import numpy as np
import pytest

import oracle as O
from hibf_helpers import layout_graph, node_split_chunks, subtree_user_bins

pytestmark = pytest.mark.gpu

THREADS = 16


def _say(msg):
    print(msg, flush=True)


def _column_mask(bins_compared, bin_words):
    m = np.zeros(bin_words, dtype=np.uint64)
    for b in bins_compared:
        m[b // 64] |= np.uint64(1) << np.uint64(b % 64)
    return m


def _columns_equal(got, want, bin_size, bin_words, mask, rows_per_chunk=1 << 22):
    g = got.reshape(bin_size, bin_words)
    w = want.reshape(bin_size, bin_words)
    for r0 in range(0, bin_size, rows_per_chunk):
        r1 = min(bin_size, r0 + rows_per_chunk)
        if not np.array_equal(g[r0:r1] & mask, w[r0:r1] & mask):
            return False
    return True


# ------------------------------------------------------------ flat IBF ----

def test_flat_ibf_8192_bins_full_matrix_and_counts():
    from hibf_amd.device import DeviceIbf, csr, to_numpy
    from hibf_amd.synth import flat_ibf_queries
    bins, bin_size, h, per_bin, seed = 8192, 1 << 23, 3, 1 << 17, 2024
    _say("flat IBF 8192 x 2^23, h=3: device emplace of 2^30 values")
    dev = DeviceIbf(bins, bin_size, h)
    dev.emplace_synthetic(seed, 0, bins, per_bin)
    _say("flat IBF: oracle emplace of the same values")
    ref = O.Ibf(bins, bin_size, h)
    ref.emplace_synthetic(seed, 0, bins, per_bin, threads=THREADS)
    _say("flat IBF: matrix comparison")
    words = dev.download()
    assert words.size == ref.words.size == bins * bin_size // 64
    assert np.array_equal(words, ref.words)
    del words
    _say("flat IBF: 1,000 queries bulk_count / 4,096 values bulk_contains")
    values, offsets, _ = flat_ibf_queries(seed + 1, 1000, 250, bins, per_bin, 0.7, seed)
    got = to_numpy(dev.bulk_count(values, offsets, 2))
    want = ref.bulk_count_batch(values, offsets, 2, threads=THREADS)
    assert np.array_equal(got, want)
    assert (want.max(axis=1) >= 175).all()  # the source bin reaches its 175 inserted values
    cv = values[:4096]
    got_c = to_numpy(dev.bulk_contains(cv))
    assert np.array_equal(got_c, np.stack([ref.bulk_contains(int(v)) for v in cv]))


# ------------------------------------------------------- HIBF at scale ----

@pytest.fixture(scope="module")
def hibf_at_scale():
    import torch
    from hibf_amd import workloads as W
    from hibf_amd.device import DeviceHibf, hll_estimates, hll_sketches_synthetic
    spec = W.HIBF_AT_SCALE
    sizes = W.hibf_user_bin_sizes(spec["size_seed"], spec["n_user_bins"])
    _say("HIBF at scale: sketches and stored layout")
    regs = hll_sketches_synthetic(sizes, spec["build_seed"], 12)
    counts = hll_estimates(regs).astype(np.uint64)
    layout = W.cached_layout(counts, spec["config"], registers=regs, write=False)
    del regs
    _say("HIBF at scale: device build")
    dev = DeviceHibf.build_synthetic(sizes, spec["build_seed"], layout_text=layout, **spec["config"])
    t = dev.tables()
    assert len(t["geoms"]) == 934 and t["depth"] == 3
    _say("HIBF at scale: arena download")
    arena = dev.download_arena()
    geoms = [O.geom(b, s, h) for (b, s, h) in t["geoms"]]
    ref = O.Hibf(geoms, arena, t["word_offsets"], t["tb_offsets"], t["next_ibf"], t["user_bin"], t["n_user_bins"])
    torch.cuda.synchronize()
    assert dev.layout_text() == layout
    return dict(dev=dev, t=t, sizes=sizes, seed=spec["build_seed"], arena=arena, ref=ref, layout=layout)


def _ibf_paths(t):
    """Technical-bin path from the root of every IBF of the exported tables."""
    paths, stack = {0: ()}, [0]
    while stack:
        i = stack.pop()
        lo = int(t["tb_offsets"][i])
        for b in range(int(t["geoms"][i][0])):
            if int(t["user_bin"][lo + b]) == O.BIN_MERGED:
                j = int(t["next_ibf"][lo + b])
                paths[j] = paths[i] + (b,)
                stack.append(j)
    return paths


def _node_of(graph, path):
    node = graph
    for b in path:
        node = next(c for c in node.children if c.parent_bin_index == b)
    return node


def _replay_values(graph, path, sizes):
    """Values the reference's k-mer set of the IBF at `path` takes in up to its last split user bin."""
    node = _node_of(graph, path)

    def subtree(n):
        return sum(int(sizes[u]) for u, _, _ in n.records) + sum(subtree(c) for c in n.children)
    last = max([i for i, r in enumerate(node.records) if r[2] > 1], default=-1)
    if last < 0 and not (node.fav is None and node.records and node.records[0][2] > 1):
        return 0
    v = 3 * subtree(node.children[node.fav]) if node.fav is not None else int(sizes[node.records[0][0]])
    return v + sum(int(sizes[u]) for u, _, _ in node.records[:last + 1])


def _synth_index(seed, u, n, keys):
    from hibf_amd.synth import synth_values
    vals = synth_values(seed, u, np.arange(int(n), dtype=np.uint64))
    # keys are the set's iteration order of exactly these values: match the two sorted orders
    assert keys.size == vals.size
    out = np.empty(keys.size, dtype=np.int64)
    out[np.argsort(keys)] = np.argsort(vals)
    return out


def _compare_ibf_columns(s, i, rng, max_merged=None, split_chunks=None):
    """Columns of IBF i against the oracle's emplace of the same values:
    merged bins, single-bin user bins and deleted bins, and, with split_chunks
    (chunk numbers of each split user bin's values, as the reference's build
    assigns them), the split user bins' technical bins."""
    from hibf_amd.synth import synth_values
    t, sizes = s["t"], s["sizes"]
    bins, bin_size, h = (int(x) for x in t["geoms"][i])
    lo = int(t["tb_offsets"][i])
    ub = [int(x) for x in t["user_bin"][lo:lo + bins]]
    nx = [int(x) for x in t["next_ibf"][lo:lo + bins]]
    merged = [b for b in range(bins) if ub[b] == O.BIN_MERGED]
    if max_merged is not None and len(merged) > max_merged:
        merged = sorted(rng.choice(merged, size=max_merged, replace=False).tolist())
    items, compared, split_runs = [], [], []
    for b in merged:
        for u in subtree_user_bins(t, nx[b]):
            items.append((u, 0, int(sizes[u]), b))
        compared.append(b)
    b = 0
    while b < bins:
        u = ub[b]
        if u == O.BIN_MERGED:
            b += 1
            continue
        if u == O.BIN_DELETED:
            compared.append(b)
            b += 1
            continue
        e = b
        while e < bins and ub[e] == u:
            e += 1
        if e - b == 1:
            items.append((u, 0, int(sizes[u]), b))
            compared.append(b)
        elif split_chunks is not None and u in split_chunks:
            split_runs.append((u, b, e))
            compared.extend(range(b, e))
        b = e
    own = O.Ibf(bins, bin_size, h)
    if items:
        a = np.asarray(items, dtype=np.uint64)
        own.emplace_synthetic_items(s["seed"], a[:, 0], a[:, 1], a[:, 2], a[:, 3], threads=THREADS)
    for u, b, e in split_runs:
        c = split_chunks[u]
        assert c.size == int(sizes[u]) and int(c.max()) == e - b - 1
        own.emplace_parallel(synth_values(s["seed"], u, np.arange(int(sizes[u]), dtype=np.uint64)),
                             np.uint64(b) + c.astype(np.uint64), threads=THREADS)
    w0 = int(t["word_offsets"][i])
    got = s["arena"][w0:w0 + own.words.size]
    bw = (bins + 63) // 64
    return _columns_equal(got, own.words, bin_size, bw, _column_mask(compared, bw)), len(compared), len(split_runs)


def _split_chunks_of(s, path, max_splits=None):
    sizes, seed = s["sizes"], s["seed"]
    return node_split_chunks(s["layout"], path, lambda st, u: st.insert_synth(seed, u, 0, int(sizes[u])),
                             lambda u, keys: _synth_index(seed, u, sizes[u], keys), max_splits)


def test_hibf_at_scale_top_level_columns(hibf_at_scale):
    """12 merged bins, every single-bin user bin and the first three split user
    bins (in the reference's build order) of the top-level IBF.  Its split
    bins come after its fullest bin, a merged bin: their chunks follow the
    reference's k-mer set after it held that bin's whole subtree (5.5e7
    values) and the split bins before them (1.9e7-2.3e7 values each)."""
    rng = np.random.default_rng(3)
    _say("HIBF at scale: replay of the top level's k-mer set (favourite subtree, first split user bins)")
    chunks = _split_chunks_of(hibf_at_scale, (), max_splits=3)
    _say(f"HIBF at scale: top-level columns (12 merged bins, single-bin and {len(chunks)} split user bins)")
    ok, n, n_split = _compare_ibf_columns(hibf_at_scale, 0, rng, max_merged=12, split_chunks=chunks)
    assert n >= 12 and n_split == len(chunks) == 3 and ok


def test_hibf_at_scale_sampled_lower_ibfs(hibf_at_scale):
    """24 sampled lower IBFs whose split user bins the oracle replays in
    seconds (at most 2e7 values through their k-mer set), all of their
    columns, split bins included."""
    rng = np.random.default_rng(4)
    s = hibf_at_scale
    paths = _ibf_paths(s["t"])
    graph = layout_graph(s["layout"])
    cheap = [i for i in range(1, len(s["t"]["geoms"])) if _replay_values(graph, paths[i], s["sizes"]) <= 20_000_000]
    pick = sorted(rng.choice(cheap, size=24, replace=False).tolist())
    _say(f"HIBF at scale: columns of IBFs {pick}")
    n_split = 0
    for i in pick:
        chunks = _split_chunks_of(s, paths[i])
        ok, _, k = _compare_ibf_columns(s, i, rng, max_merged=16, split_chunks=chunks)
        n_split += k
        assert ok, i
    assert n_split >= 10


def test_hibf_at_scale_membership_and_count(hibf_at_scale):
    from hibf_amd import workloads as W
    from hibf_amd.device import to_numpy
    s = hibf_at_scale
    dev, ref = s["dev"], s["ref"]
    lengths = W.query_lengths(40, 600, fixed=250)
    v, o, _ = W.hibf_queries(41, s["sizes"], s["seed"], lengths, 0.8)
    thr = 175  # uint16_t(0.7 * 250), as test/tutorial/search.cpp:42 computes it
    _say("HIBF at scale: membership_for of 600 queries")
    ids, hoff = dev.membership_for(v, o, thr)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    want_ids, want_off = ref.membership_batch(v, o, thr)
    assert np.array_equal(hoff, want_off) and np.array_equal(ids, want_ids)
    assert (np.diff(hoff.astype(np.int64)) >= 1).mean() > 0.95
    _say("HIBF at scale: bulk_count of 120 queries")
    o2 = o[:121]
    got = to_numpy(dev.bulk_count(v[: int(o2[-1])], o2, thr, 2))
    assert np.array_equal(got, ref.bulk_count_batch(v[: int(o2[-1])], o2, thr, 2, threads=THREADS))


def test_batched_form_one_graph_batch_of_65536(hibf_at_scale):
    import torch
    from hibf_amd import workloads as W
    from hibf_amd.device import Agent, as_device_u64, to_numpy
    s = hibf_at_scale
    dev, ref = s["dev"], s["ref"]
    G = 65536
    lengths = W.query_lengths(42, G)
    v_h, o_h, _ = W.hibf_queries(43, s["sizes"], s["seed"], lengths, 0.6)
    thr_h = np.minimum((0.7 * np.diff(o_h.astype(np.int64))).astype(np.int64), 65535).astype(np.uint16)
    dv, do = as_device_u64(v_h), as_device_u64(o_h)
    dt = torch.from_numpy(thr_h.view(np.int16)).cuda().view(torch.uint16)
    cap = 8 * G
    agent = Agent()
    dev.membership_prepare(agent, G)
    ids = torch.empty(cap, dtype=torch.uint64, device="cuda")
    hs = torch.empty(G, dtype=torch.uint64, device="cuda")
    hc = torch.empty(G, dtype=torch.uint64, device="cuda")
    a = (agent, dv, do, dt, ids, hs, hc)
    dev.membership_launch(*a)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        dev.membership_launch(*a)
    torch.cuda.synchronize()

    _say("batched form: 65,536 plain launches")
    agent1 = Agent()
    dev.membership_prepare(agent1, 1)
    per = 64
    ids1 = torch.zeros(G * per, dtype=torch.uint64, device="cuda")
    hs1 = torch.zeros(G, dtype=torch.uint64, device="cuda")
    hc1 = torch.zeros(G, dtype=torch.uint64, device="cuda")
    for q in range(G):
        dev.membership_launch(agent1, dv, do[q:q + 2], dt[q:q + 1], ids1[q * per:(q + 1) * per], hs1[q:q + 1],
                              hc1[q:q + 1])
    torch.cuda.synchronize()
    assert not agent1.status()[1]
    ids1_n, hs1_n, hc1_n = to_numpy(ids1), to_numpy(hs1), to_numpy(hc1)
    plain = [ids1_n[q * per + int(hs1_n[q]): q * per + int(hs1_n[q]) + int(hc1_n[q])].tolist() for q in range(G)]

    _say("batched form: oracle on all 65,536 queries")
    want_ids, want_off = ref.membership_batch(v_h, o_h, thr_h)
    want = [want_ids[int(want_off[q]):int(want_off[q + 1])].tolist() for q in range(G)]
    assert plain == want

    for replay in range(2):
        _say(f"batched form: graph replay {replay}")
        ids.fill_(0)
        g.replay()
        torch.cuda.synchronize()
        assert not agent.status()[1]
        ids_n, hs_n, hc_n = to_numpy(ids), to_numpy(hs), to_numpy(hc)
        got = [ids_n[int(hs_n[q]):int(hs_n[q]) + int(hc_n[q])].tolist() for q in range(G)]
        assert got == want, replay
        dev.membership_for(v_h, o_h[:9], thr_h[:8])  # eager work between replays
        torch.cuda.synchronize()


def test_serving_500k_qps_mix_alone_oracle(hibf_at_scale):
    import torch
    from hibf_amd import workloads as W
    from hibf_amd.device import to_numpy
    s = hibf_at_scale
    dev, ref = s["dev"], s["ref"]
    rate, seconds = 500_000, 0.2
    nq = int(rate * seconds)
    lengths = W.query_lengths(44, nq)
    v_h, o_h, _ = W.hibf_queries(45, s["sizes"], s["seed"], lengths, 0.6)
    thr_h = np.minimum((0.7 * np.diff(o_h.astype(np.int64))).astype(np.int64), 65535).astype(np.uint16)
    rng = np.random.default_rng(46)
    arrival = np.cumsum(rng.exponential(1.0 / rate, size=nq))
    arrival -= arrival[0]
    v_in = torch.empty(v_h.size, dtype=torch.int64, pin_memory=True)
    v_in.numpy()[:] = v_h.view(np.int64)
    R = 16
    _say(f"serving: {nq} requests at {rate} q/s, window 200 us")
    lat, ids, cnt, st = dev.serve(v_in, o_h, thr_h, arrival, window_us=200, streams=4, budget_bytes=4 << 30,
                                  results_per_query=R)
    assert (lat > 0).all() and st["batches"] > 100
    _say("serving: batched answer of every request")
    b_ids, b_off = dev.membership_for(v_h, o_h, thr_h)
    b_ids, b_off = to_numpy(b_ids), to_numpy(b_off)
    for q in range(nq):
        want = b_ids[int(b_off[q]):int(b_off[q + 1])]
        assert int(cnt[q]) == want.size and ids[q, :min(R, want.size)].tolist() == want[:R].tolist(), q
    _say("serving: 64 requests alone, 500 against the oracle")
    for q in rng.choice(nq, size=64, replace=False):
        q = int(q)
        a_ids, a_off = dev.membership_for(v_h[int(o_h[q]):int(o_h[q + 1])], np.array([0, o_h[q + 1] - o_h[q]],
                                                                                      dtype=np.uint64),
                                          thr_h[q:q + 1])
        alone = to_numpy(a_ids).tolist()
        assert int(cnt[q]) == len(alone) and ids[q, :min(R, len(alone))].tolist() == alone[:R]
    for q in rng.choice(nq, size=500, replace=False):
        q = int(q)
        want = ref.membership_for(v_h[int(o_h[q]):int(o_h[q + 1])], int(thr_h[q]))
        assert int(cnt[q]) == len(want) and ids[q, :min(R, len(want))].tolist() == want[:R]
