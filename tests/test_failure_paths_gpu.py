"""Failure paths forced from the host, every kernel still exiting normally:
a serving request over the admission budget (HIBF_GPU_ERR_BUDGET), malformed
serving parameters, device allocations refused by the library's allocation
cap (HIBF_GPU_ERR_OUT_OF_MEMORY, then the same calls succeeding once the cap
is lifted), device buffers of the wrong size rejected by the Python
wrappers, and a synchronous membership call whose first attempt overflowed
its temporary id buffer leaving no dropped flag behind."""
import numpy as np
import pytest

import oracle as O
from hibf_amd import native
from hibf_helpers import oracle_of

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _user_bins(rng, n, lo, hi):
    return [rng.integers(0, 2**63, size=int(k), dtype=np.int64).astype(np.uint64) for k in rng.integers(lo, hi, size=n)]


def _queries(rng, ubs, nq, length):
    qs = []
    for _ in range(nq):
        ub = ubs[int(rng.integers(0, len(ubs)))]
        k = int(0.8 * length)
        qs.append(np.concatenate([ub[rng.integers(0, ub.size, size=k)],
                                  rng.integers(0, 2**63, size=length - k, dtype=np.int64).astype(np.uint64)]))
    return qs


@pytest.fixture(scope="module")
def small_hibf():
    from hibf_amd.device import DeviceHibf, csr
    rng = np.random.default_rng(91)
    ubs = _user_bins(rng, 200, 50, 500)
    dev = DeviceHibf.build(ubs, tmax=64)
    qs = _queries(rng, ubs, 300, 120)
    values, offsets = csr(qs)
    thr = np.full(len(qs), 60, dtype=np.uint16)
    arrival = np.cumsum(rng.exponential(1e-5, size=len(qs)))
    arrival -= arrival[0]
    return dev, ubs, qs, values, offsets, thr, arrival


def _serve_code(dev, values, offsets, thr, arrival, **kw):
    try:
        dev.serve(values, offsets, thr, arrival, **kw)
    except native.HibfGpuError as e:
        return e.code, native.lib().hibf_gpu_last_error().decode()
    return native.OK, ""


def test_serving_request_over_budget(small_hibf):
    dev, _, _, values, offsets, thr, arrival = small_hibf
    _say("serving over budget")
    code, msg = _serve_code(dev, values, offsets, thr, arrival, budget_bytes=512)
    assert code == native.ERR_BUDGET and "more than the budget" in msg
    # one byte short of the largest request's footprint (its values, offset, threshold and 16 result ids)
    need = int(np.diff(offsets.astype(np.int64)).max()) * 8 + 34 + 8 * 16
    assert _serve_code(dev, values, offsets, thr, arrival, budget_bytes=need - 1)[0] == native.ERR_BUDGET
    assert _serve_code(dev, values, offsets, thr, arrival, budget_bytes=need)[0] == native.OK


@pytest.mark.parametrize("kw, message", [
    (dict(streams=0), "must be > 0"),
    (dict(results_per_query=0), "must be > 0"),
    (dict(max_batch=0), "must be > 0"),
    (dict(max_batch_values=0), "must be > 0"),
    (dict(max_batch_values=100), "larger than max_batch_values"),
    (dict(streams=65), "at most 64 streams"),
    (dict(window_us=-1.0), "window_us"),
    (dict(window_us=float("nan")), "window_us"),
])
def test_serving_rejects_malformed_parameters(small_hibf, kw, message):
    dev, _, _, values, offsets, thr, arrival = small_hibf
    _say(f"serving malformed parameters {kw}")
    code, msg = _serve_code(dev, values, offsets, thr, arrival, **kw)
    assert code == native.ERR_INVALID_ARGUMENT and message in msg, msg


def test_serving_rejects_malformed_requests(small_hibf):
    dev, _, _, values, offsets, thr, arrival = small_hibf
    _say("serving malformed requests")
    back = arrival.copy()
    back[5], back[6] = back[6] + 1.0, back[5]
    code, msg = _serve_code(dev, values, offsets, thr, back)
    assert code == native.ERR_INVALID_ARGUMENT and "non-decreasing" in msg
    with pytest.raises(ValueError):
        dev.serve(values, offsets, thr[:-1], arrival)
    with pytest.raises(ValueError):
        dev.serve(values, offsets, thr, arrival[:-1])
    with pytest.raises(ValueError):
        dev.serve(values[:-1], offsets, thr, arrival)


def test_allocation_cap_forces_out_of_memory_then_recovers():
    from hibf_amd.device import Agent, DeviceHibf, DeviceIbf, csr, to_numpy
    rng = np.random.default_rng(92)
    ubs = _user_bins(rng, 40, 20_000, 40_000)
    small = DeviceIbf(64, 4096, 2)  # 32 KiB: made before the cap
    vals = rng.integers(0, 2**63, size=5000, dtype=np.int64).astype(np.uint64)
    small.emplace(vals, rng.integers(0, 64, size=vals.size).astype(np.uint64))
    qs = [vals[rng.integers(0, vals.size, 100)] for _ in range(3000)]
    values, offsets = csr(qs)
    try:
        native.tune_set("alloc.max_bytes", 64 << 10)
        assert native.tune_get("alloc.max_bytes") == 64 << 10
        _say("allocation cap: IBF, HIBF, agent workspace, serving")
        with pytest.raises(native.HibfGpuError) as e:
            DeviceIbf(4096, 1 << 16, 2)  # 32 MiB bit matrix
        assert e.value.code == native.ERR_OUT_OF_MEMORY
        with pytest.raises(native.HibfGpuError) as e:
            DeviceHibf.build(ubs, tmax=64)
        assert e.value.code == native.ERR_OUT_OF_MEMORY
        agent = Agent()
        with pytest.raises(native.HibfGpuError) as e:
            small.membership_for(values, offsets, 50, agent)  # uint16 counts of 3,000 queries: 384 KiB
        assert e.value.code == native.ERR_OUT_OF_MEMORY
    finally:
        native.tune_set("alloc.max_bytes", 0)
    _say("allocation cap lifted: the same calls succeed")
    ref = O.Ibf.from_words(64, 4096, 2, small.download())
    ids, hoff = small.membership_for(values, offsets, 50, agent)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    for q in range(0, len(qs), 97):
        assert ids[int(hoff[q]):int(hoff[q + 1])].tolist() == ref.membership_for(qs[q], 50)
    dev = DeviceHibf.build(ubs, tmax=64)
    oh = oracle_of(dev)
    q = np.concatenate([ubs[3][:200], rng.integers(0, 2**63, size=50, dtype=np.int64).astype(np.uint64)])
    got, off = dev.membership_for(q, np.array([0, q.size], dtype=np.uint64), 150)
    assert to_numpy(got).tolist() == oh.membership_for(q, 150)


def test_device_buffers_of_the_wrong_size_are_rejected(small_hibf):
    import torch
    from hibf_amd.device import Agent, DeviceIbf, as_device_u64
    dev, _, _, values, offsets, thr, _ = small_hibf
    _say("buffer size checks")
    d = torch.device("cuda", 0)
    v, o = as_device_u64(values), as_device_u64(offsets)
    nq = o.numel() - 1
    t16 = torch.zeros(nq, dtype=torch.uint16, device=d)
    ids = torch.zeros(1000, dtype=torch.uint64, device=d)
    hs = torch.zeros(nq, dtype=torch.uint64, device=d)
    agent = Agent()
    dev.membership_prepare(agent, nq)
    for bad in (dict(thresholds=t16[:-1]), dict(hit_start=hs[:-1]), dict(hit_count=hs[:-1]),
                dict(thresholds=torch.zeros(nq, dtype=torch.uint64, device=d)), dict(values=values)):
        a = dict(values=v, offsets=o, thresholds=t16, ids=ids, hit_start=hs, hit_count=hs.clone())
        a.update(bad)
        with pytest.raises(ValueError):
            dev.membership_launch(agent, a["values"], a["offsets"], a["thresholds"], a["ids"], a["hit_start"],
                                  a["hit_count"])
    t64 = torch.zeros(nq, dtype=torch.uint64, device=d)
    out = torch.zeros((nq, dev.number_of_user_bins), dtype=torch.uint16, device=d)
    with pytest.raises(ValueError):
        dev.bulk_count_into(agent, v, o, t64[:-1], out, 2)
    with pytest.raises(ValueError):
        dev.bulk_count_into(agent, v, o, t64, out[:-1], 2)
    with pytest.raises(ValueError):
        dev.bulk_count_into(agent, v, o, t64, out, 4)
    ibf = DeviceIbf(100, 1024, 2)
    cnt = torch.zeros((nq, 100), dtype=torch.uint16, device=d)
    with pytest.raises(ValueError):
        ibf.bulk_count_into(v, o, cnt[:-1], 2)
    with pytest.raises(ValueError):
        ibf.bulk_count_into(v, o, cnt, 1)
    ibf.bulk_count_into(v, o, cnt, 2)  # the right sizes pass
    torch.cuda.synchronize()


def test_sync_membership_retry_leaves_no_dropped_flag():
    from hibf_amd.device import Agent, DeviceHibf, csr, to_numpy
    rng = np.random.default_rng(93)
    ubs = _user_bins(rng, 2000, 20, 60)
    dev = DeviceHibf.build(ubs, tmax=64)
    oh = oracle_of(dev)
    # threshold 0: every user bin is reported for every query, 2,000 ids each, far past the
    # first attempt's temporary buffer (max(queries, 1024) ids)
    qs = [ubs[i][:10] for i in range(8)]
    values, offsets = csr(qs)
    agent = Agent()
    _say("sync membership with a retry")
    ids, hoff = dev.membership_for(values, offsets, 0, agent)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    assert int(hoff[-1]) == 8 * 2000
    for q in range(len(qs)):
        assert ids[int(hoff[q]):int(hoff[q + 1])].tolist() == oh.membership_for(qs[q], 0)
    total, dropped = agent.status()
    assert total == 8 * 2000 and not dropped
