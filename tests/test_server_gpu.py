"""The online serving loop (hibf_gpu_server_*): requests submitted from other
host threads while earlier batches run on the device; every result equals
the request's answer alone, the batched answer and the oracle's."""
import threading
import time

import numpy as np
import pytest

from hibf_amd import native
from hibf_helpers import oracle_of

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def served():
    from hibf_amd.device import DeviceHibf, csr, to_numpy
    rng = np.random.default_rng(31)
    ubs = [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
           for n in rng.integers(50, 2000, size=500)]
    dev = DeviceHibf.build(ubs, tmax=64)
    qs, ts = [], []
    for i in range(4000):
        L = int(rng.integers(20, 600))
        src = ubs[int(rng.integers(0, len(ubs)))]
        k = int(0.7 * L)
        qs.append(np.concatenate([src[rng.integers(0, src.size, k)],
                                  rng.integers(0, 2**63, L - k, dtype=np.int64).astype(np.uint64)]))
        ts.append(1 if i % 97 == 0 else int(0.5 * L))  # some requests with many ids
    values, offsets = csr(qs)
    ids, hoff = dev.membership_for(values, offsets, np.array(ts, dtype=np.uint16))
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    batched = [ids[int(hoff[q]):int(hoff[q + 1])].tolist() for q in range(len(qs))]
    return dev, qs, ts, batched, oracle_of(dev)


def _collect(server, n, deadline_s=120.0):
    got = {}
    t_end = time.monotonic() + deadline_s
    while len(got) < n and time.monotonic() < t_end:
        for ticket, ids, count, lat in server.poll(1024, wait_s=0.05):
            got[ticket] = (ids, count, lat)
    return got


def test_submissions_from_two_threads_while_batches_run(served):
    from hibf_amd.device import Server
    dev, qs, ts, batched, ref = served
    R = 16
    server = Server(dev, window_us=300, streams=3, max_batch=256, results_per_query=R)
    tickets = {}
    lock = threading.Lock()
    errors = []

    def submitter(part, seed):
        rng = np.random.default_rng(seed)
        try:
            for q in part:
                t = server.submit(qs[q], ts[q])
                with lock:
                    tickets[t] = q
                time.sleep(float(rng.exponential(20e-6)))  # Poisson arrivals, about 50k/s per thread
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    _say("server: two submitting threads, polling on the main thread")
    order = np.random.default_rng(5).permutation(len(qs))
    threads = [threading.Thread(target=submitter, args=(order[i::2], i)) for i in range(2)]
    for th in threads:
        th.start()
    got = _collect(server, len(qs))
    for th in threads:
        th.join()
    assert not errors
    server.drain(30.0)
    got.update({t: (ids, c, lat) for t, ids, c, lat in server.poll(len(qs))})
    st = server.stats()
    server.close()
    assert len(got) == len(qs) == len(tickets)
    assert st["batches"] > 1
    for ticket, (ids, count, lat) in got.items():
        q = tickets[ticket]
        want = batched[q]
        assert count == len(want) and ids == want[:R], q
        assert lat > 0
    for q in range(0, len(qs), 41):
        assert batched[q] == ref.membership_for(qs[q], ts[q])


@pytest.mark.parametrize("wave", [0, 1])
def test_each_request_keeps_its_own_ids_when_one_batch_overflows(served, wave):
    """A batch whose requests find far more ids than batch x R in total (every
    97th request has threshold 1 and hundreds of ids) still gives every
    request its own first R ids: a trace replayed as one batch, and the same
    requests submitted to the online server within one window; with the
    query-per-block descent and with the query-per-wave descent."""
    native.tune_set("hibf.wave_kernel", wave)
    try:
        _overflow_case(served)
    finally:
        native.tune_set("hibf.wave_kernel", 0)


def _overflow_case(served):
    from hibf_amd.device import Server, csr
    dev, qs, ts, batched, _ = served
    R = 4
    sel = list(range(0, 97 * 8, 97)) + list(range(1, 200))  # 8 many-id requests first, then 199 others
    assert sum(len(batched[q]) for q in sel) > 2 * len(sel) * R
    _say("serve: one batch whose ids overflow batch x R")
    values, offsets = csr([qs[q] for q in sel])
    lat, ids, cnt, st = dev.serve(values, offsets, np.array([ts[q] for q in sel], dtype=np.uint16),
                                  np.zeros(len(sel)), window_us=1e5, streams=1, results_per_query=R)
    assert st["batches"] == 1
    for i, q in enumerate(sel):
        want = batched[q]
        assert int(cnt[i]) == len(want), q
        assert ids[i, :min(len(want), R)].tolist() == want[:R], q
    _say("server: the same requests within one window")
    server = Server(dev, window_us=2e5, streams=1, results_per_query=R)
    tickets = {server.submit(qs[q], ts[q]): q for q in sel}
    got = _collect(server, len(sel))
    st = server.stats()
    server.close()
    assert len(got) == len(sel)
    assert st["batches"] <= 2
    for ticket, (got_ids, count, _) in got.items():
        q = tickets[ticket]
        assert count == len(batched[q]) and got_ids == batched[q][:R], q


def test_requests_alone_equal_requests_in_the_mix(served):
    from hibf_amd.device import Server
    dev, qs, ts, batched, _ = served
    _say("server: requests one at a time, each waited for")
    server = Server(dev, window_us=0, streams=2, results_per_query=8)
    for q in range(0, len(qs), 157):
        t = server.submit(qs[q], ts[q])
        res = server.poll(4, wait_s=10.0)
        assert len(res) == 1 and res[0][0] == t
        assert res[0][2] == len(batched[q]) and res[0][1] == batched[q][:8]
    server.close()


def test_server_rejects_what_it_cannot_admit(served):
    from hibf_amd.device import Server
    dev, qs, ts, _, _ = served
    _say("server: budget and size limits")
    server = Server(dev, budget_bytes=4096, max_batch_values=300, results_per_query=4)
    long_q = np.arange(1000, dtype=np.uint64)
    with pytest.raises(native.HibfGpuError) as e:
        server.submit(long_q, 1)
    assert e.value.code == native.ERR_INVALID_ARGUMENT
    server.close()
    server = Server(dev, budget_bytes=2048, results_per_query=4)
    with pytest.raises(native.HibfGpuError) as e:
        server.submit(np.arange(300, dtype=np.uint64), 1)  # 300 x 8 bytes + 66 > 2,048
    assert e.value.code == native.ERR_BUDGET
    t = server.submit(np.arange(100, dtype=np.uint64), 1)  # fits: answered
    res = server.poll(4, wait_s=10.0)
    assert len(res) == 1 and res[0][0] == t
    server.close()
    with pytest.raises(native.HibfGpuError):
        native.check(native.lib().hibf_gpu_server_create(dev.h, None, None))
