"""GPU parity on the benchmark configurations that fit a test: the small plain
HIBF (64 user bins of 5k-20k values, 10,000 queries x 200 values, threshold
140, plus bulk_count), and the batched form (log-normal query lengths in
hipGraph-captured batches) against the plain one-launch-per-query form."""
import numpy as np
import pytest

from hibf_helpers import oracle_of

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _distinct_u64(rng, n):
    v = np.unique(rng.integers(0, 2**63, size=n + 64, dtype=np.int64).astype(np.uint64) * np.uint64(2)
                  + rng.integers(0, 2, size=n + 64).astype(np.uint64))
    rng.shuffle(v)
    return v[:n]


def _queries(rng, ubs, nq, length, frac_in):
    qs = []
    for _ in range(nq):
        ub = ubs[int(rng.integers(0, len(ubs)))]
        k = int(round(frac_in * length))
        q = np.concatenate([ub[rng.integers(0, ub.size, size=k)],
                            rng.integers(0, 2**63, size=length - k, dtype=np.int64).astype(np.uint64)])
        qs.append(q)
    return qs


def test_plain_hibf_64_user_bins_full_size():
    """64 user bins of 5,000-20,000 distinct uniform uint64 values, h = 2,
    FPR 0.05; 10,000 queries x 200 values (80 % from one user bin); every
    query's membership list (t = 140) and uint16 counts (t = 1 and 140)
    equal the oracle's."""
    from hibf_amd.device import DeviceHibf, csr, to_numpy
    _say("plain HIBF 64 user bins, 10,000 x 200")
    rng = np.random.default_rng(64)
    ubs = [_distinct_u64(rng, int(n)) for n in rng.integers(5000, 20001, size=64)]
    dev = DeviceHibf.build(ubs, number_of_hash_functions=2, maximum_fpr=0.05, threads=1)
    ref = oracle_of(dev)
    qs = _queries(rng, ubs, 10_000, 200, 0.8)
    values, offsets = csr(qs)
    ids, hoff = dev.membership_for(values, offsets, 140)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    want_ids, want_off = ref.membership_batch(values, offsets, 140)
    assert np.array_equal(hoff, want_off)
    assert np.array_equal(ids, want_ids)
    assert (np.diff(hoff.astype(np.int64)) >= 1).mean() > 0.95  # the source user bin is found
    for t in (1, 140):
        got = to_numpy(dev.bulk_count(values, offsets, t, 2))
        assert np.array_equal(got, ref.bulk_count_batch(values, offsets, t, 2, threads=8)), t


def test_graph_batches_equal_plain_per_query_launches():
    """Log-normal query lengths (median 300, [50, 5000], 60 % from one user
    bin, t = 0.7 L) on a 2,000-user-bin HIBF: hipGraph batches, replayed three
    times with eager work in between, give the same ids as one launch per
    query and as the oracle."""
    import torch
    from hibf_amd import workloads as W
    from hibf_amd.device import Agent, DeviceHibf, as_device_u64, to_numpy
    _say("graph batches vs plain launches, 2,000 user bins")
    rng = np.random.default_rng(2000)
    sizes = np.clip(np.round(rng.lognormal(np.log(2000), 1.0, size=2000)), 100, 50_000).astype(np.uint64)
    dev = DeviceHibf.build_synthetic(sizes, 3, tmax=64)
    ref = oracle_of(dev)
    nq = 1500
    lengths = W.query_lengths(11, nq)
    v_h, o_h, _ = W.hibf_queries(12, sizes, 3, lengths, 0.6)
    thr_h = np.minimum((0.7 * np.diff(o_h.astype(np.int64))).astype(np.int64), 65535).astype(np.uint16)
    dv, do = as_device_u64(v_h), as_device_u64(o_h)
    dt = torch.from_numpy(thr_h.view(np.int16)).cuda().view(torch.uint16)

    G = 512
    nb = (nq + G - 1) // G
    cap = 64 * G
    agent = Agent()
    dev.membership_prepare(agent, G)
    ids_b = torch.empty(nb * cap, dtype=torch.uint64, device="cuda")
    hs = torch.empty(nq, dtype=torch.uint64, device="cuda")
    hc = torch.empty(nq, dtype=torch.uint64, device="cuda")
    graphs = []
    for b in range(nb):
        q0, q1 = b * G, min(nq, (b + 1) * G)
        a = (agent, dv, do[q0:q1 + 1], dt[q0:q1], ids_b[b * cap:(b + 1) * cap], hs[q0:q1], hc[q0:q1])
        dev.membership_launch(*a)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            dev.membership_launch(*a)
        graphs.append(g)
    torch.cuda.synchronize()

    def batched_lists():
        ids_n, hs_n, hc_n = to_numpy(ids_b), to_numpy(hs), to_numpy(hc)
        out = []
        for q in range(nq):
            s0 = (q // G) * cap + int(hs_n[q])
            out.append(ids_n[s0:s0 + int(hc_n[q])].tolist())
        return out

    # Plain form: one launch per query.
    agent1 = Agent()
    dev.membership_prepare(agent1, 1)
    ids1 = torch.empty(1 << 14, dtype=torch.uint64, device="cuda")
    hs1 = torch.empty(1, dtype=torch.uint64, device="cuda")
    hc1 = torch.empty(1, dtype=torch.uint64, device="cuda")
    plain = []
    for q in range(nq):
        dev.membership_launch(agent1, dv, do[q:q + 2], dt[q:q + 1], ids1, hs1, hc1)
        torch.cuda.synchronize()
        s0, n = int(to_numpy(hs1)[0]), int(to_numpy(hc1)[0])
        plain.append(to_numpy(ids1[s0:s0 + n]).tolist())
    for q in range(0, nq, 7):
        assert plain[q] == ref.membership_for(v_h[int(o_h[q]):int(o_h[q + 1])], int(thr_h[q])), q

    for replay in range(3):
        ids_b.fill_(0)
        for g in graphs:
            g.replay()
        torch.cuda.synchronize()
        _, dropped = agent.status()
        assert not dropped
        assert batched_lists() == plain, replay
        # eager work between replays
        dev.membership_for(v_h, o_h[:5], thr_h[:4])
        torch.cuda.synchronize()


def test_graph_batches_over_three_streams_equal_one_stream():
    """The batched form as bench.py runs it: graph batches alternating over
    three streams, one agent per stream, so consecutive batches overlap on the
    device.  Three rounds of replays (with eager work between them) give the
    per-query lists of the same batches replayed on one stream, and the
    oracle's on a sample."""
    import torch
    from hibf_amd import workloads as W
    from hibf_amd.device import Agent, DeviceHibf, as_device_u64, to_numpy
    _say("graph batches over three streams vs one stream, 2,000 user bins")
    rng = np.random.default_rng(2001)
    sizes = np.clip(np.round(rng.lognormal(np.log(2000), 1.0, size=2000)), 100, 50_000).astype(np.uint64)
    dev = DeviceHibf.build_synthetic(sizes, 5, tmax=64)
    ref = oracle_of(dev)
    nq = 6000
    lengths = W.query_lengths(13, nq)
    v_h, o_h, _ = W.hibf_queries(14, sizes, 5, lengths, 0.6)
    thr_h = np.minimum((0.7 * np.diff(o_h.astype(np.int64))).astype(np.int64), 65535).astype(np.uint16)
    dv, do = as_device_u64(v_h), as_device_u64(o_h)
    dt = torch.from_numpy(thr_h.view(np.int16)).cuda().view(torch.uint16)
    G, NS = 500, 3
    nb = (nq + G - 1) // G
    cap = 64 * G

    def capture(agents):
        ids_b = torch.empty(nb * cap, dtype=torch.uint64, device="cuda")
        hs = torch.empty(nq, dtype=torch.uint64, device="cuda")
        hc = torch.empty(nq, dtype=torch.uint64, device="cuda")
        for ag in agents:
            dev.membership_prepare(ag, G)
        graphs = []
        for b in range(nb):
            q0, q1 = b * G, min(nq, (b + 1) * G)
            a = (agents[b % len(agents)], dv, do[q0:q1 + 1], dt[q0:q1], ids_b[b * cap:(b + 1) * cap], hs[q0:q1],
                 hc[q0:q1])
            dev.membership_launch(*a)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                dev.membership_launch(*a)
            graphs.append(g)
        torch.cuda.synchronize()
        return graphs, ids_b, hs, hc

    def lists(ids_b, hs, hc):
        ids_n, hs_n, hc_n = to_numpy(ids_b), to_numpy(hs), to_numpy(hc)
        return [ids_n[(q // G) * cap + int(hs_n[q]):(q // G) * cap + int(hs_n[q]) + int(hc_n[q])].tolist()
                for q in range(nq)]

    one_agent = Agent()
    g1, ids1, hs1, hc1 = capture([one_agent])
    for g in g1:
        g.replay()
    torch.cuda.synchronize()
    serial = lists(ids1, hs1, hc1)
    for q in range(0, nq, 11):
        assert serial[q] == ref.membership_for(v_h[int(o_h[q]):int(o_h[q + 1])], int(thr_h[q])), q

    agents = [Agent() for _ in range(NS)]
    streams = [torch.cuda.current_stream()] + [torch.cuda.Stream() for _ in range(NS - 1)]
    g3, ids3, hs3, hc3 = capture(agents)
    for rnd in range(3):
        ids3.fill_(0)
        hc3.fill_(0)
        start = torch.cuda.Event()
        start.record(streams[0])
        for sx in streams[1:]:
            sx.wait_event(start)
        for b, g in enumerate(g3):
            with torch.cuda.stream(streams[b % NS]):
                g.replay()
        for sx in streams[1:]:
            streams[0].wait_stream(sx)
        torch.cuda.synchronize()
        for ag in agents:
            assert not ag.status()[1]
        assert lists(ids3, hs3, hc3) == serial, rnd
        dev.membership_for(v_h, o_h[:5], thr_h[:4])  # eager work between rounds
        torch.cuda.synchronize()
