"""GPU parity of the HyperLogLog sketches the build uses for user-bin
cardinalities (src/sketch/hyperloglog.cpp, compute_sketches.cpp): registers
byte for byte and estimates bit for bit against the oracle, merges, and the
reference's own known answers."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SMALL = [255881241332063154, 13132817195163223578, 5120631300412165844, 16862690419523416066,
         148518882728022940, 15892358469365346306, 10885195586503739779, 9563173945158404745]


def _say(msg):
    print(msg, flush=True)


def _sets(rng, sizes):
    return [rng.integers(0, 2**63, size=int(n), dtype=np.int64).astype(np.uint64) * np.uint64(2) +
            rng.integers(0, 2, size=int(n)).astype(np.uint64) for n in sizes]


def test_hll_reference_known_answers():
    from hibf_amd.device import hll_estimates, hll_sketches, to_numpy
    _say("hll known answers")
    regs = hll_sketches([np.asarray(SMALL, dtype=np.uint64), np.zeros(0, dtype=np.uint64)], bits=5)
    est = hll_estimates(regs)
    assert abs(est[0] - 7.899522493) < 1e-7 and est[1] == 0.0
    assert np.array_equal(to_numpy(regs)[0], O.hll_sketch(SMALL, 5))
    vals = O.mt19937_64(0, 1500)
    assert hll_estimates(hll_sketches([vals], bits=5))[0] == O.hll_estimate(O.hll_sketch(vals, 5))


# 65,536 values is one chunk of the device kernel; larger sets are merged from several.
@pytest.mark.parametrize("bits", [5, 6, 7, 12, 16, 17, 20])
def test_hll_sketches_and_estimates_match_oracle(bits):
    from hibf_amd.device import hll_estimates, hll_sketches, to_numpy
    _say(f"hll bits={bits}")
    rng = np.random.default_rng(bits)
    sizes = [0, 1, 7, 100, 5000, 65535, 65536, 65537, 200000, 3]
    sets = _sets(rng, sizes)
    sets[3] = np.concatenate([sets[3], sets[3]])  # duplicates do not change a sketch
    regs_d = hll_sketches(sets, bits=bits)
    est = hll_estimates(regs_d)
    regs = to_numpy(regs_d)
    for i, s in enumerate(sets):
        want = O.hll_sketch(s, bits)
        assert np.array_equal(regs[i], want), (bits, i)
        assert est[i] == O.hll_estimate(want), (bits, i)


def test_hll_merge_matches_oracle():
    from hibf_amd.device import hll_estimates, hll_merge, hll_sketches, to_numpy
    _say("hll merge")
    rng = np.random.default_rng(5)
    sets = _sets(rng, rng.integers(0, 3000, size=40))
    regs = hll_sketches(sets, bits=12)
    groups = [[0], [1, 2, 3], list(range(40)), [], [7, 7, 39]]
    merged = hll_merge(regs, groups)
    got = to_numpy(merged)
    host = to_numpy(regs)
    for g, row in zip(groups, got):
        want = np.zeros(4096, dtype=np.uint8)
        for k in g:
            want = O.hll_merge(want, host[k])
        assert np.array_equal(row, want)
    # merge == sketch of the union (hyperloglog_test.cpp:130-163)
    union = np.concatenate([sets[1], sets[2], sets[3]])
    assert np.array_equal(got[1], O.hll_sketch(union, 12))
    assert hll_estimates(merged)[1] == O.hll_estimate(O.hll_sketch(union, 12))


def test_hll_bit_width_is_checked():
    from hibf_amd import native
    from hibf_amd.device import hll_sketches
    _say("hll bit width")
    with pytest.raises(native.HibfGpuError, match=r"bit width must be in the range \[5,32\]"):
        hll_sketches([np.arange(10, dtype=np.uint64)], bits=4)
    # 33 bits: refused before any pointer is read
    assert native.lib().hibf_gpu_hll_sketch(None, None, 1, 33, None, None) == native.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("bits,npos", [(5, 9), (7, 65), (10, 130), (12, 257), (12, 2)])
def test_union_estimates_match_oracle(bits, npos):
    """All union estimates of the layout DP (every column j and k < j of a
    shuffled order, empty and repeated sketches included) equal the oracle's
    sequential merge_and_estimate, bit for bit after truncation; the first
    column entries are the prefix unions."""
    from hibf_amd.device import hll_sketches, hll_union_estimates, to_numpy
    _say(f"union estimates bits={bits} npos={npos}")
    rng = np.random.default_rng(bits * 1000 + npos)
    n = npos + 3
    sizes = rng.integers(0, 4000, size=n)
    sizes[1] = 0
    sets = _sets(rng, sizes)
    sets[2] = sets[4]
    regs = hll_sketches(sets, bits=bits)
    pos = rng.permutation(n)[:npos]
    got = to_numpy(hll_union_estimates(regs, pos))
    want = O.hll_union_estimates(to_numpy(regs), pos)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_union_estimates_refuse_wide_sketches():
    from hibf_amd.device import hll_sketches, hll_union_estimates
    from hibf_amd.native import HibfGpuError
    regs = hll_sketches([np.arange(10, dtype=np.uint64)] * 3, bits=13)
    with pytest.raises(HibfGpuError):
        hll_union_estimates(regs, [0, 1, 2])


def _host_sketched_layout(tmp_path, counts, regs, bits, tmax):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "layout_union"
    r = subprocess.run(["g++", "-std=c++20", "-O2", f"-I{root}/include", f"-I{root}/hibf_amd/csrc",
                        os.path.join(root, "tests", "cpp", "layout_union.cpp"),
                        os.path.join(root, "hibf_amd", "csrc", "layout.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    f = tmp_path / "sketches.bin"
    f.write_bytes(np.asarray(counts, dtype=np.uint64).tobytes() + regs.tobytes())
    r = subprocess.run([str(exe), "sketched", str(f), str(len(counts)), str(bits), str(tmax)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("n,bits,tmax", [(200, 12, 64), (700, 10, 64), (1500, 12, 128)])
def test_sketched_layout_on_device_equals_host(tmp_path, n, bits, tmax):
    """compute_layout with sketches (union estimates + similarity
    rearrangement): the device path (triangle of union estimates, pair
    estimates of the clustering) gives the layout text of the host path, which
    merges sketches one by one as the reference (its toolbox pinned by the
    reference's known answers)."""
    from hibf_amd import native
    from hibf_amd.device import hll_estimates, hll_sketches, to_numpy
    _say(f"sketched layout n={n} bits={bits} tmax={tmax}")
    rng = np.random.default_rng(n)
    sizes = np.clip(rng.lognormal(np.log(2000), 1.2, size=n), 50, 40000).astype(np.int64)
    sets = _sets(rng, sizes)
    for i in range(0, n - 3, 7):  # overlapping user bins give the clustering real choices
        sets[i + 1] = np.concatenate([sets[i + 1], sets[i][: len(sets[i]) // 2]])
    regs = hll_sketches(sets, bits=bits)
    counts = hll_estimates(regs).astype(np.uint64)
    c = native.validate_config(native.make_config(number_of_user_bins=n, tmax=tmax, sketch_bits=bits))
    got = native.compute_layout_sketched(c, counts, regs.data_ptr())
    want = _host_sketched_layout(tmp_path, counts, to_numpy(regs), bits, tmax)
    assert got == want
    plain = native.compute_layout(c, counts)
    assert plain != got  # union estimates and rearrangement changed the layout


def test_wide_sketch_build_uses_host_union_estimates():
    """Sketches of 13 bits are past the device union estimates: the layout's
    union and pair estimates run on the host (the reference's sequential
    merges), give the layout of compute_layout_sketched's host path, and the
    built HIBF answers like the oracle on its tables."""
    from hibf_amd import native
    from hibf_amd.device import DeviceHibf, csr, hll_estimates, hll_sketches, to_numpy
    from hibf_helpers import oracle_of
    _say("13-bit sketches")
    rng = np.random.default_rng(13)
    ubs = _sets(rng, rng.integers(50, 800, size=90))
    dev = DeviceHibf.build(ubs, tmax=64, sketch_bits=13)
    regs = hll_sketches([np.unique(u) for u in ubs], bits=13)
    c = native.validate_config(native.make_config(number_of_user_bins=len(ubs), tmax=64, sketch_bits=13))
    text = native.compute_layout_sketched(c, hll_estimates(regs).astype(np.uint64), regs.data_ptr())
    same = DeviceHibf.build(ubs, layout_text=text, tmax=64, sketch_bits=13)
    assert np.array_equal(dev.download_arena(), same.download_arena())
    ref = oracle_of(dev)
    qs = [np.concatenate([ubs[int(i)][:30], _sets(rng, [10])[0]]) for i in rng.integers(0, len(ubs), size=40)]
    v, o = csr(qs)
    ids, hoff = dev.membership_for(v, o, 25)
    ids, hoff = to_numpy(ids), to_numpy(hoff)
    for q in range(len(qs)):
        assert ids[int(hoff[q]):int(hoff[q + 1])].tolist() == ref.membership_for(qs[q], 25)
