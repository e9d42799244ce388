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
