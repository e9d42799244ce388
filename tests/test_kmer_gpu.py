"""GPU parity of the reads-in path: device k-mer hashes equal the oracle's
restatement, and the tutorial search (k-mers of each read, threshold
uint16_t(0.7 * #k-mers)) returns the oracle's ids."""
import random

import numpy as np
import pytest

import oracle as O
from hibf_helpers import oracle_of

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _reads(rng, n, lo, hi, alphabet="ACGT"):
    return ["".join(rng.choice(alphabet) for _ in range(rng.randint(lo, hi))) for _ in range(n)]


@pytest.mark.parametrize("k", [1, 12, 19, 31, 32])
def test_kmer_hash_matches_oracle(k):
    from hibf_amd.device import kmer_hash, to_numpy
    _say(f"kmer hash k={k}")
    rng = random.Random(k)
    reads = _reads(rng, 300, 0, 400, "ACGTACGTACGTacgtNU") + ["", "A" * (k - 1), "C" * k]
    # Reads longer than the kernel's 8,192-base segments, and lengths at the segment seams.
    reads += _reads(rng, 1, 20000, 20000) + _reads(rng, 1, 8192, 8192) + _reads(rng, 1, 8192 + k - 1, 8192 + k - 1)
    reads += _reads(rng, 1, 8192 + k, 8192 + k) + _reads(rng, 1, 16384 - k + 2, 16384 - k + 2)
    hashes, offsets = kmer_hash(reads, k)
    h, o = to_numpy(hashes), to_numpy(offsets)
    assert o[0] == 0 and o.size == len(reads) + 1
    for i, r in enumerate(reads):
        assert h[int(o[i]):int(o[i + 1])].tolist() == O.kmer_hashes(r, k).tolist(), i


def test_kmer_hash_rejects_bad_k():
    from hibf_amd.device import kmer_hash
    from hibf_amd.native import HibfGpuError
    for k in (0, 33):
        with pytest.raises(HibfGpuError):
            kmer_hash(["ACGT"], k)


_MINIMISER_CASES = [
    # (shape, window_size, seed)
    (0b1111, 4, 0),                          # window of one k-mer: every canonical value
    ((1 << 19) - 1, 23, O.MINIMISER_SEED),   # ungapped k = 19, 5 k-mers per window
    ((1 << 32) - 1, 40, O.MINIMISER_SEED),   # k = 32 fills 64 bits
    ((1 << 12) - 1, 70, 0),                  # 59 k-mers per window
    (0b1101, 8, 0),                          # gapped
    (0b11011001110111, 30, O.MINIMISER_SEED),  # gapped, size 14, count 10
    ((1 << 16) - 1, 16 + 4095, 12345),       # the largest window: 4,096 k-mers
]


@pytest.mark.parametrize("shape,window,seed", _MINIMISER_CASES)
def test_minimiser_hash_matches_oracle(shape, window, seed):
    from hibf_amd.device import minimiser_hash, to_numpy
    _say(f"minimiser hash shape={shape:#x} window={window} seed={seed:#x}")
    rng = random.Random(shape ^ window)
    size = shape.bit_length()
    w = window - size + 1
    reads = _reads(rng, 200, 0, 300, "ACGTACGTACGTacgtNU") + ["", "A" * (size - 1), "C" * size, "A" * 500]
    # Low-complexity reads (long runs of equal values: ties everywhere) and
    # reads past the kernel's 512-position segments, at and around the seams.
    reads += ["ACG" * 300, "AAAAAAAAAAAAAAAAAAAACAAAAAAAAAAAAAAAAAAAA" * 30]
    for nk in (512, 513, 511 + w, 512 + w, 1024 + w - 1, 3000):
        reads += _reads(rng, 1, nk + size - 1, nk + size - 1)
    reads += _reads(rng, 1, 9000, 9000)
    hashes, offsets = minimiser_hash(reads, shape, window, seed)
    h, o = to_numpy(hashes), to_numpy(offsets)
    assert o[0] == 0 and o.size == len(reads) + 1
    for i, r in enumerate(reads):
        assert h[int(o[i]):int(o[i + 1])].tolist() == O.minimiser_hashes(r, shape, window, seed).tolist(), i


def test_minimiser_hash_rejects_bad_arguments():
    from hibf_amd.device import minimiser_hash
    from hibf_amd.native import HibfGpuError
    for shape, window in ((0, 8), (0b1111, 3), ((1 << 33) - 1, 40), (0b1111, 4 + 4096)):
        with pytest.raises(HibfGpuError):
            minimiser_hash(["ACGTACGTAC"], shape, window)


def test_search_reads_like_the_tutorial():
    """User bins are k-mer sets of random genomes; reads are sampled from them
    (some with errors, some random); ids equal the oracle's membership_for on
    the oracle's hashes with threshold uint16_t(0.7 * #k-mers)."""
    from hibf_amd.device import DeviceHibf
    _say("tutorial search on reads")
    k = 19
    rng = random.Random(7)
    genomes = ["".join(rng.choice("ACGT") for _ in range(rng.randint(2000, 6000))) for _ in range(120)]
    ubs = [np.unique(O.kmer_hashes(g, k)) for g in genomes]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    reads = []
    for _ in range(400):
        g = genomes[rng.randrange(len(genomes))]
        start = rng.randrange(0, len(g) - 250)
        read = list(g[start:start + rng.randint(100, 250)])
        for _ in range(rng.randint(0, 3)):
            read[rng.randrange(len(read))] = rng.choice("ACGTN")
        reads.append("".join(read))
    reads += _reads(rng, 50, 0, 200) + ["", "ACGT"]
    got = dev.membership_for_reads(reads, k, 0.7)
    found = 0
    for r, g in zip(reads, got):
        hashes = O.kmer_hashes(r, k)
        assert g == ref.membership_for(hashes, O.kmer_threshold(hashes.size, 0.7))
        found += bool(g)
    assert found >= 380


@pytest.mark.parametrize("pinned", [True, False], ids=["pinned", "pageable"])
def test_serving_reads_equals_the_tutorial_search(pinned):
    """hibf_gpu_hibf_serve_reads (bases copied, hashed and thresholded on the
    device inside the serving loop) answers every read as the oracle's
    membership_for on its k-mer hashes with threshold uint16_t(0.7 * #k-mers),
    for Poisson arrivals and small batching windows; empty and short reads
    included."""
    from hibf_amd.device import DeviceHibf
    _say(f"serving reads pinned={pinned}")
    k = 19
    rng = random.Random(11)
    genomes = ["".join(rng.choice("ACGT") for _ in range(rng.randint(2000, 5000))) for _ in range(80)]
    ubs = [np.unique(O.kmer_hashes(g, k)) for g in genomes]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    reads = []
    for _ in range(3000):
        g = genomes[rng.randrange(len(genomes))]
        start = rng.randrange(0, len(g) - 250)
        reads.append(g[start:start + rng.randint(50, 250)])
    reads[5], reads[17], reads[100] = "", "ACG", "N" * 30
    nrng = np.random.default_rng(3)
    arrival = np.cumsum(nrng.exponential(1.0 / 200000.0, size=len(reads)))
    for window in (50.0, 500.0):
        lat, ids, cnt, st = dev.serve_reads(reads, arrival, k=k, fraction=0.7, window_us=window, max_batch=512,
                                            max_batch_bases=1 << 18, results_per_query=16, pinned=pinned)
        assert st["batches"] >= 2 and (lat > 0).all()
        for i, r in enumerate(reads):
            hashes = O.kmer_hashes(r, k)
            want = ref.membership_for(hashes, O.kmer_threshold(hashes.size, 0.7))
            assert int(cnt[i]) == len(want), i
            assert ids[i, :min(len(want), 16)].tolist() == want[:16], i


def test_search_reads_by_minimisers():
    """User bins are the minimiser sets of random genomes; reads sampled from
    them (some with errors) are queried by their minimisers with threshold
    uint16_t(0.7 * #minimisers); ids equal the oracle's membership_for on the
    oracle's minimisers."""
    from hibf_amd.device import DeviceHibf
    _say("search by minimisers")
    shape, window = (1 << 15) - 1, 25
    rng = random.Random(9)
    genomes = ["".join(rng.choice("ACGT") for _ in range(rng.randint(2000, 6000))) for _ in range(80)]
    ubs = [np.unique(O.minimiser_hashes(g, shape, window)) for g in genomes]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    reads = []
    for _ in range(300):
        g = rng.choice(genomes)
        at = rng.randint(0, len(g) - 150)
        r = list(g[at:at + 150])
        for _ in range(rng.randint(0, 3)):
            r[rng.randrange(150)] = rng.choice("ACGT")
        reads.append("".join(r))
    reads += ["".join(rng.choice("ACGT") for _ in range(150)) for _ in range(20)] + ["ACG", ""]
    got = dev.membership_for_reads(reads, 0, 0.7, minimiser=(shape, window))
    for i, r in enumerate(reads):
        hashes = O.minimiser_hashes(r, shape, window)
        assert got[i] == ref.membership_for(hashes, O.kmer_threshold(hashes.size, 0.7)), i


def test_serving_reads_by_minimisers():
    """hibf_gpu_hibf_serve_reads_minimisers (bases copied, minimiser counts,
    offsets, thresholds and hashes computed on the device inside the serving
    loop) answers every read as the oracle's membership_for on its minimisers
    with threshold uint16_t(0.7 * #minimisers), for Poisson arrivals and two
    batching windows; empty, short and N-only reads included; a window shorter
    than the shape is rejected."""
    from hibf_amd.device import DeviceHibf
    from hibf_amd.native import HibfGpuError
    _say("serving reads by minimisers")
    shape, window = 0b11011011101111111, 24   # gapped, size 17
    rng = random.Random(21)
    genomes = ["".join(rng.choice("ACGT") for _ in range(rng.randint(2000, 5000))) for _ in range(60)]
    ubs = [np.unique(O.minimiser_hashes(g, shape, window)) for g in genomes]
    dev = DeviceHibf.build(ubs, tmax=64)
    ref = oracle_of(dev)
    reads = []
    for _ in range(2000):
        g = genomes[rng.randrange(len(genomes))]
        start = rng.randrange(0, len(g) - 300)
        reads.append(g[start:start + rng.randint(40, 300)])
    reads[3], reads[9], reads[50] = "", "ACGTACGT", "N" * 40
    nrng = np.random.default_rng(5)
    arrival = np.cumsum(nrng.exponential(1.0 / 200000.0, size=len(reads)))
    for win_us in (50.0, 500.0):
        lat, ids, cnt, st = dev.serve_reads(reads, arrival, fraction=0.7, window_us=win_us, max_batch=512,
                                            max_batch_bases=1 << 17, results_per_query=16,
                                            minimiser=(shape, window))
        assert st["batches"] >= 2 and (lat > 0).all()
        for i, r in enumerate(reads):
            hashes = O.minimiser_hashes(r, shape, window)
            want = ref.membership_for(hashes, O.kmer_threshold(hashes.size, 0.7))
            assert int(cnt[i]) == len(want), i
            assert ids[i, :min(len(want), 16)].tolist() == want[:16], i
    with pytest.raises(HibfGpuError):
        dev.serve_reads(reads[:10], arrival[:10], minimiser=(shape, 10))
