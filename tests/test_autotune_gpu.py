"""GPU tests of the launch-shape autotuner: every candidate must reproduce the
default shape's output (the tuner raises otherwise), the result equals the
oracle, and a saved record applies again."""
import json

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture
def restore_knobs():
    from hibf_amd import autotune, native
    keys = list(autotune.FLAT_SPACE) + list(autotune.HIBF_SPACE)
    saved = {k: native.tune_get(k) for k in keys}
    yield
    for k, v in saved.items():
        native.tune_set(k, v)


def test_tune_flat_count(tmp_path, restore_knobs):
    import torch
    from hibf_amd import autotune, native
    from hibf_amd.device import DeviceIbf, as_device_u64, csr, to_numpy
    _say("autotune flat bulk_count 8192 bins")
    rng = np.random.default_rng(5)
    vals = rng.integers(0, 2**63, size=40000, dtype=np.int64).astype(np.uint64)
    bidx = rng.integers(0, 8192, size=vals.size).astype(np.uint64)
    ref = O.Ibf(8192, 2048, 3)
    ref.emplace(vals, bidx)
    dev = DeviceIbf(8192, 2048, 3)
    dev.emplace(vals, bidx)
    qs = [vals[rng.integers(0, vals.size, size=L)] for L in rng.integers(0, 400, size=300)]
    v_h, o_h = csr(qs)
    rec = autotune.tune_flat_count(dev, as_device_u64(v_h), as_device_u64(o_h), 2, reps=1)
    assert len(rec["candidates"]) == len(autotune._candidates(autotune.FLAT_SPACE))
    assert {k: native.tune_get(k) for k in autotune.FLAT_SPACE} == rec["best"]
    got = to_numpy(dev.bulk_count(v_h, o_h, 2))
    assert np.array_equal(got, ref.bulk_count_batch(v_h, o_h, 2))
    path = tmp_path / "flat.json"
    autotune.save_record(rec, str(path))
    for k in autotune.FLAT_SPACE:
        native.tune_set(k, autotune.FLAT_SPACE[k][0])
    assert autotune.apply_record(autotune.load_record(str(path))) == rec["best"]
    assert json.loads(path.read_text())["workload"]["bins"] == 8192
    torch.cuda.synchronize()


def test_tune_hibf_membership(restore_knobs):
    from hibf_amd import autotune
    from hibf_amd.device import DeviceHibf, csr, to_numpy
    from hibf_helpers import oracle_of
    _say("autotune hibf membership 600 user bins")
    rng = np.random.default_rng(9)
    ubs = [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
           for n in rng.integers(20, 300, size=600)]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    qs = [np.concatenate([ubs[int(rng.integers(0, 600))][:40], rng.integers(0, 2**63, size=20, dtype=np.int64)
                          .astype(np.uint64)]) for _ in range(200)]
    v_h, o_h = csr(qs)
    thr = np.full(len(qs), 30, dtype=np.uint16)
    import torch
    from hibf_amd.device import as_device_u64
    t = torch.from_numpy(thr.view(np.int16)).cuda().view(torch.uint16)
    rec = autotune.tune_hibf_membership(dev, as_device_u64(v_h), as_device_u64(o_h), t, reps=1)
    assert len(rec["candidates"]) == len(autotune._candidates(autotune.HIBF_SPACE))
    ids, hoff = dev.membership_for(v_h, o_h, thr)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    for q in range(len(qs)):
        assert ids[hoff[q]:hoff[q + 1]].tolist() == ref.membership_for(qs[q], 30)


def test_tune_hibf_membership_launch_form(restore_knobs):
    """The launch-form tuner (the form a hipGraph batch replays): every
    candidate's per-query lists equal the default shape's (the tuner raises
    otherwise), and the lists of the chosen shape equal the oracle's."""
    import torch
    from hibf_amd import autotune, native
    from hibf_amd.device import Agent, DeviceHibf, as_device_u64, csr, to_numpy
    from hibf_helpers import oracle_of
    _say("autotune hibf membership, launch form, 700 user bins")
    rng = np.random.default_rng(19)
    ubs = [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
           for n in rng.integers(20, 400, size=700)]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    qs = []
    for _ in range(3000):
        L = int(rng.integers(0, 300))
        src = ubs[int(rng.integers(0, 700))]
        k = int(0.8 * L)
        qs.append(np.concatenate([src[rng.integers(0, src.size, k)],
                                  rng.integers(0, 2**63, size=L - k, dtype=np.int64).astype(np.uint64)]))
    v_h, o_h = csr(qs)
    thr = np.array([max(1, int(0.7 * q.size)) for q in qs], dtype=np.uint16)
    thr[::50] = 0  # some queries report every user bin
    nq = len(qs)
    t = torch.from_numpy(thr.view(np.int16)).cuda().view(torch.uint16)
    ids = torch.empty(8 * nq + 64 * 700, dtype=torch.uint64, device="cuda")
    hs = torch.empty(nq, dtype=torch.uint64, device="cuda")
    hc = torch.empty(nq, dtype=torch.uint64, device="cuda")
    agent = Agent()
    rec = autotune.tune_hibf_membership_launch(dev, agent, as_device_u64(v_h), as_device_u64(o_h), t, ids, hs, hc,
                                               reps=1)
    assert len(rec["candidates"]) == len(autotune._candidates(autotune.HIBF_SPACE))
    assert {k: native.tune_get(k) for k in autotune.HIBF_SPACE} == rec["best"]
    dev.membership_prepare(agent, nq)
    dev.membership_launch(agent, as_device_u64(v_h), as_device_u64(o_h), t, ids, hs, hc)
    torch.cuda.synchronize()
    total, dropped = agent.status()
    assert not dropped
    ids_n, hs_n, hc_n = to_numpy(ids), to_numpy(hs), to_numpy(hc)
    assert int(hc_n.sum()) == total
    for q in range(0, nq, 7):
        got = ids_n[int(hs_n[q]):int(hs_n[q]) + int(hc_n[q])].tolist()
        assert got == ref.membership_for(qs[q], int(thr[q])), q


def test_tune_hibf_membership_batches_over_streams(restore_knobs):
    """The overlapped-batches tuner: 3,000 queries in batches of 500 over three
    streams with one agent each; every candidate's per-query lists equal the
    default shape's, and the chosen shape's lists equal the oracle's."""
    import torch
    from hibf_amd import autotune, native
    from hibf_amd.device import Agent, DeviceHibf, as_device_u64, csr, to_numpy
    from hibf_helpers import oracle_of
    _say("autotune hibf membership, batches over three streams, 700 user bins")
    rng = np.random.default_rng(23)
    ubs = [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64)
           for n in rng.integers(20, 400, size=700)]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    qs = []
    for _ in range(3000):
        L = int(rng.integers(0, 300))
        src = ubs[int(rng.integers(0, 700))]
        k = int(0.8 * L)
        qs.append(np.concatenate([src[rng.integers(0, src.size, k)],
                                  rng.integers(0, 2**63, size=L - k, dtype=np.int64).astype(np.uint64)]))
    v_h, o_h = csr(qs)
    thr = np.array([max(1, int(0.7 * q.size)) for q in qs], dtype=np.uint16)
    nq, batch = len(qs), 500
    t = torch.from_numpy(thr.view(np.int16)).cuda().view(torch.uint16)
    ids = torch.empty(6 * (8 * batch + 64 * 700), dtype=torch.uint64, device="cuda")
    hs = torch.empty(nq, dtype=torch.uint64, device="cuda")
    hc = torch.empty(nq, dtype=torch.uint64, device="cuda")
    agents = [Agent() for _ in range(3)]
    streams = [torch.cuda.current_stream()] + [torch.cuda.Stream() for _ in range(2)]
    rec = autotune.tune_hibf_membership_batches(dev, agents, streams, as_device_u64(v_h), as_device_u64(o_h), t,
                                                ids, hs, hc, batch, reps=1)
    assert len(rec["candidates"]) == len(autotune._candidates(autotune.HIBF_SPACE))
    assert {k: native.tune_get(k) for k in autotune.HIBF_SPACE} == rec["best"]
    torch.cuda.synchronize()
    for ag in agents:
        assert not ag.status()[1]
    ids_n, hs_n, hc_n = to_numpy(ids), to_numpy(hs), to_numpy(hc)
    cap = ids.numel() // 6  # batch b's hit_start values count from ids[b * cap]
    for q in range(0, nq, 7):
        at = q // batch * cap + int(hs_n[q])
        got = ids_n[at:at + int(hc_n[q])].tolist()
        assert got == ref.membership_for(qs[q], int(thr[q])), q
