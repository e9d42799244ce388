"""GPU tests of the IBF/HIBF archives (cereal binary layout): what the library
writes parses field by field under the layout rules, what the layout rules
write loads into HBM word for word, loaded filters answer as the originals and
the oracle do, and archives embed in wrapper files at byte offsets."""
import numpy as np
import pytest

import cereal_layout as CL
import oracle as O
from hibf_helpers import oracle_of

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rand(rng, n):
    return rng.integers(0, 2**63, size=n, dtype=np.int64).astype(np.uint64)


def _ibf_pair(bins, bin_size, h, seed):
    from hibf_amd.device import DeviceIbf
    rng = np.random.default_rng(seed)
    vals = _rand(rng, 4000)
    bidx = rng.integers(0, bins, size=vals.size).astype(np.uint64)
    ref = O.Ibf(bins, bin_size, h)
    ref.emplace(vals, bidx)
    dev = DeviceIbf(bins, bin_size, h)
    dev.emplace(vals, bidx)
    return ref, dev, vals, rng


@pytest.mark.parametrize("bins,bin_size,h", [(3, 1000, 2), (130, 777, 3), (8192, 300, 2)])
def test_ibf_save_parses_and_loads(tmp_path, bins, bin_size, h):
    from hibf_amd.device import DeviceIbf, csr, to_numpy
    _say(f"ibf archive bins={bins} bin_size={bin_size} h={h}")
    ref, dev, vals, rng = _ibf_pair(bins, bin_size, h, bins)
    occ = rng.integers(0, 1000, size=int(dev.geometry.technical_bins)).astype(np.uint64)
    dev.set_occupancy(occ, True)
    p = tmp_path / "ibf.bin"
    dev.save(p)
    rec, end = CL.parse_ibf(p.read_bytes())
    assert end == p.stat().st_size
    g = dev.geometry
    assert (rec.version, rec.bins, rec.technical_bins, rec.bin_size, rec.hash_shift, rec.bin_words, rec.hash_funs,
            rec.bits) == (1, bins, g.technical_bins, bin_size, g.hash_shift, g.bin_words, h, g.technical_bins * bin_size)
    assert np.array_equal(rec.words, ref.words)
    assert np.array_equal(rec.occupancy, occ) and rec.track_occupancy
    back, end2 = DeviceIbf.load(p)
    assert end2 == end
    assert np.array_equal(back.download(), ref.words)
    occ2, track2 = back.occupancy()
    assert np.array_equal(occ2, occ) and track2
    qs = [np.concatenate([vals[rng.integers(0, vals.size, size=150)], _rand(rng, 50)]) for _ in range(40)]
    v, o = csr(qs)
    assert np.array_equal(to_numpy(back.bulk_count(v, o, 2)), ref.bulk_count_batch(v, o, 2))


def test_ibf_written_by_layout_rules_loads(tmp_path):
    from hibf_amd.device import DeviceIbf, to_numpy
    _say("ibf archive from the layout rules")
    ref, _, vals, rng = _ibf_pair(200, 513, 2, 5)
    occ = np.arange(256, dtype=np.uint64)
    p = tmp_path / "ibf.bin"
    p.write_bytes(CL.ibf_bytes(CL.ibf_record(200, 513, 2, ref.words, occ, False)))
    dev, end = DeviceIbf.load(p)
    assert end == p.stat().st_size
    assert np.array_equal(dev.download(), ref.words)
    assert dev.occupancy()[0].tolist() == occ.tolist() and not dev.occupancy()[1]
    got = to_numpy(dev.bulk_contains(vals[:50]))
    assert np.array_equal(got, np.stack([ref.bulk_contains(int(v)) for v in vals[:50]]))


def test_archives_in_a_wrapper_file(tmp_path):
    """A caller's format (a k-mer byte, then two IBFs, then an HIBF), appended
    piece by piece and read back at byte offsets."""
    from hibf_amd.device import DeviceHibf, DeviceIbf
    _say("archives at offsets")
    ref_a, a, _, _ = _ibf_pair(70, 100, 2, 1)
    ref_b, b, _, _ = _ibf_pair(64, 333, 3, 2)
    rng = np.random.default_rng(3)
    hibf = DeviceHibf.build([_rand(rng, int(n)) for n in rng.integers(10, 100, size=90)], tmax=64)
    p = tmp_path / "index.bin"
    p.write_bytes(b"\x13")
    a.save(p, append=True)
    b.save(p, append=True)
    hibf.save(p, append=True)
    a2, off = DeviceIbf.load(p, 1)
    b2, off = DeviceIbf.load(p, off)
    h2, off = DeviceHibf.load(p, off)
    assert off == p.stat().st_size
    assert np.array_equal(a2.download(), ref_a.words) and np.array_equal(b2.download(), ref_b.words)
    assert np.array_equal(h2.download_arena(), hibf.download_arena())


@pytest.mark.parametrize("n_ub,tmax", [(300, 64), (1500, 128)])
def test_hibf_save_parses_and_loads(tmp_path, n_ub, tmax):
    from hibf_amd.device import DeviceHibf, csr, to_numpy
    _say(f"hibf archive n_ub={n_ub} tmax={tmax}")
    rng = np.random.default_rng(n_ub)
    ubs = [_rand(rng, int(n)) for n in rng.integers(20, 400, size=n_ub)]
    dev = DeviceHibf.build(ubs, tmax=tmax)
    tab = dev.tables()
    arena = dev.download_arena()
    p = tmp_path / "hibf.bin"
    dev.save(p)
    rec, end = CL.parse_hibf(p.read_bytes())
    assert end == p.stat().st_size and rec.version == 1 and rec.number_of_user_bins == n_ub
    n = len(tab["geoms"])
    assert len(rec.ibfs) == n
    to, nx, ub, wo = tab["tb_offsets"], tab["next_ibf"], tab["user_bin"], tab["word_offsets"]
    expect_prev = [(0, 0)] * n
    for i in range(n):
        bins, bin_size, h = tab["geoms"][i]
        r = rec.ibfs[i]
        assert (r.bins, r.bin_size, r.hash_funs, r.track_occupancy) == (bins, bin_size, h, False)
        assert np.array_equal(r.words, arena[int(wo[i]):int(wo[i]) + r.words.size])
        assert not r.occupancy.any() and r.occupancy.size == r.technical_bins
        assert np.array_equal(rec.next_ibf_id[i], nx[int(to[i]):int(to[i + 1])])
        assert np.array_equal(rec.ibf_bin_to_user_bin_id[i], ub[int(to[i]):int(to[i + 1])])
        for b in range(bins):
            if int(ub[int(to[i]) + b]) == 2**64 - 1:
                expect_prev[int(nx[int(to[i]) + b])] = (i, b)
    assert [tuple(x) for x in rec.prev_ibf_id] == expect_prev

    back, end2 = DeviceHibf.load(p)
    assert end2 == end
    ref = oracle_of(back)
    qs = [np.concatenate([ubs[int(rng.integers(0, n_ub))][:30], _rand(rng, 10)]) for _ in range(300)]
    v, o = csr(qs)
    ids0, off0 = dev.membership_for(v, o, 20)
    ids1, off1 = back.membership_for(v, o, 20)
    assert np.array_equal(to_numpy(ids0), to_numpy(ids1)) and np.array_equal(to_numpy(off0), to_numpy(off1))
    ids1, off1 = to_numpy(ids1), to_numpy(off1)
    for q in range(0, len(qs), 10):
        assert ids1[int(off1[q]):int(off1[q + 1])].tolist() == ref.membership_for(qs[q], 20)
    # Saving the loaded HIBF gives the same bytes.
    p2 = tmp_path / "hibf2.bin"
    back.save(p2)
    assert p2.read_bytes() == p.read_bytes()


def test_hibf_written_by_layout_rules_loads(tmp_path):
    """An HIBF archive assembled by the layout rules from a built HIBF's
    tables, with occupancy and track flags set, loads and answers alike."""
    from hibf_amd.device import DeviceHibf, csr, to_numpy
    _say("hibf archive from the layout rules")
    rng = np.random.default_rng(8)
    ubs = [_rand(rng, int(n)) for n in rng.integers(20, 300, size=200)]
    dev = DeviceHibf.build(ubs, tmax=64)
    tab = dev.tables()
    arena = dev.download_arena()
    to, wo = tab["tb_offsets"], tab["word_offsets"]
    ibfs, nxt, ubt = [], [], []
    for i, (bins, bin_size, h) in enumerate(tab["geoms"]):
        n_words = ((bins + 63) // 64) * bin_size
        occ = rng.integers(0, 50, size=((bins + 63) // 64) * 64).astype(np.uint64)
        ibfs.append(CL.ibf_record(bins, bin_size, h, arena[int(wo[i]):int(wo[i]) + n_words], occ, True))
        nxt.append(tab["next_ibf"][int(to[i]):int(to[i + 1])])
        ubt.append(tab["user_bin"][int(to[i]):int(to[i + 1])])
    rec = CL.HibfRecord(200, ibfs, nxt, [(0, 0)] * len(ibfs), ubt)
    p = tmp_path / "hibf.bin"
    p.write_bytes(CL.hibf_bytes(rec))
    back, end = DeviceHibf.load(p)
    assert end == p.stat().st_size
    qs = [np.concatenate([ubs[int(rng.integers(0, 200))][:25], _rand(rng, 5)]) for _ in range(100)]
    v, o = csr(qs)
    a, b = dev.membership_for(v, o, 18), back.membership_for(v, o, 18)
    assert np.array_equal(to_numpy(a[0]), to_numpy(b[0])) and np.array_equal(to_numpy(a[1]), to_numpy(b[1]))
