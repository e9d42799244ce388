"""GPU test of hibf_gpu_hibf_membership_launch_noalloc, the membership launch
that allocates nothing (its workspace is a fixed array in the library): its
ids must equal those of hibf_gpu_hibf_membership_launch (agent workspace) on
the same queries, on two calls in a row, the second with fewer queries, for
both descent shapes (block per query, wave per query), and the oracle's."""
import ctypes

import numpy as np
import pytest

import oracle as O
from hibf_helpers import oracle_of

pytestmark = pytest.mark.gpu

MAX_QUERIES = 65536


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def hibf_and_queries():
    import torch
    from hibf_amd.device import DeviceHibf, csr
    rng = np.random.default_rng(77)
    ubs = [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
           for n in rng.integers(50, 3000, size=600)]
    dev = DeviceHibf.build(ubs, tmax=64)
    qs = []
    for i in range(3000):
        L = int(rng.integers(20, 400))
        src = ubs[int(rng.integers(0, len(ubs)))]
        k = int(0.8 * L)
        qs.append(np.concatenate([src[rng.integers(0, src.size, k)],
                                  rng.integers(0, 2**63, L - k, dtype=np.int64).astype(np.uint64)]))
    values, offsets = csr(qs)
    thr = np.array([int(0.6 * q.size) for q in qs], dtype=np.uint16)
    torch.cuda.synchronize()
    return dev, ubs, qs, values, offsets, thr


def _lists(ids, hs, hc, nq):
    from hibf_amd.device import to_numpy
    ids, hs, hc = to_numpy(ids), to_numpy(hs), to_numpy(hc)
    return [ids[int(hs[q]):int(hs[q]) + int(hc[q])].tolist() for q in range(nq)]


def _run(dev, values, offsets, thr, nq, noalloc, agent=None):
    import torch
    from hibf_amd import native
    from hibf_amd.device import as_device_u64
    d = torch.device("cuda", 0)
    v = as_device_u64(values[: int(offsets[nq])])
    o = as_device_u64(offsets[: nq + 1])
    t = torch.from_numpy(np.ascontiguousarray(thr[:nq]).view(np.int16)).to(d).view(torch.uint16)
    cap = 16 * nq + 64
    ids = torch.full((cap,), 7, dtype=torch.uint64, device=d)
    hs = torch.zeros(nq, dtype=torch.uint64, device=d)
    hc = torch.zeros(nq, dtype=torch.uint64, device=d)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    if noalloc:
        native.check(native.lib().hibf_gpu_hibf_membership_launch_noalloc(
            dev.h, P(v), P(o), nq, P(t), P(ids), cap, P(hs), P(hc), s))
    else:
        dev.membership_prepare(agent, nq)
        dev.membership_launch(agent, v, o, t, ids, hs, hc)
    torch.cuda.synchronize()
    got = _lists(ids, hs, hc, nq)
    assert int((hs.cpu().view(torch.int64) + hc.cpu().view(torch.int64)).max()) <= cap, "ids dropped"
    return got


@pytest.mark.parametrize("wave", [0, 1])
def test_noalloc_equals_agent_launch(hibf_and_queries, wave):
    from hibf_amd import native
    from hibf_amd.device import Agent
    dev, ubs, qs, values, offsets, thr = hibf_and_queries
    native.tune_set("hibf.wave_kernel", wave)
    try:
        agent = Agent()
        for nq in (len(qs), 700):  # two calls in a row, the second smaller
            _say(f"noalloc wave={wave} nq={nq}")
            want = _run(dev, values, offsets, thr, nq, noalloc=False, agent=agent)
            got = _run(dev, values, offsets, thr, nq, noalloc=True)
            assert got == want
        oh = oracle_of(dev)
        for q in range(0, 700, 37):
            assert got[q] == oh.membership_for(qs[q], int(thr[q]))
    finally:
        native.tune_set("hibf.wave_kernel", 0)


def test_noalloc_rejects_oversized_batch(hibf_and_queries):
    from hibf_amd import native
    dev = hibf_and_queries[0]
    rc = native.lib().hibf_gpu_hibf_membership_launch_noalloc(
        dev.h, ctypes.c_void_p(8), ctypes.c_void_p(8), MAX_QUERIES + 1, ctypes.c_void_p(8), ctypes.c_void_p(8), 1,
        ctypes.c_void_p(8), ctypes.c_void_p(8), None)
    assert rc == native.ERR_INVALID_ARGUMENT
    assert b"65536" in native.lib().hibf_gpu_last_error()
