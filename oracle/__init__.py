"""Oracle for the HIBF query path (test infrastructure only).

Python bindings to ``liboracle.so`` (``hibf_oracle.c``), a plain-C restatement
of the reference query path, plus an independent pure-Python big-integer
restatement of the hash used to cross-check it.  Only ``tests/``, the smoke
check in ``__graft_entry__.py`` and the CPU-baseline leg of ``bench.py`` import
this package, and only as the checker.  The product package ``hibf_amd`` never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

BIN_MERGED = 0xFFFFFFFFFFFFFFFF  # hierarchical_interleaved_bloom_filter.hpp:34
BIN_DELETED = 0xFFFFFFFFFFFFFFFE  # hierarchical_interleaved_bloom_filter.hpp:35

# interleaved_bloom_filter.hpp:172-176
HASH_SEEDS = (
    13572355802537770549,
    13043817825332782213,
    10650232656628343401,
    16499269484942379435,
    4893150838803335377,
)

_u64p = ctypes.POINTER(ctypes.c_uint64)


class OrcIbf(ctypes.Structure):
    """Geometry of one IBF (interleaved_bloom_filter.hpp:160-170)."""

    _fields_ = [
        ("bins", ctypes.c_uint64),
        ("technical_bins", ctypes.c_uint64),
        ("bin_size", ctypes.c_uint64),
        ("hash_shift", ctypes.c_uint64),
        ("bin_words", ctypes.c_uint64),
        ("hash_funs", ctypes.c_uint64),
    ]

    @property
    def word_count(self) -> int:
        return self.technical_bins * self.bin_size // 64


class OrcHibf(ctypes.Structure):
    _fields_ = [
        ("n_ibfs", ctypes.c_uint64),
        ("ibfs", ctypes.POINTER(OrcIbf)),
        ("arena", _u64p),
        ("word_offsets", _u64p),
        ("tb_offsets", _u64p),
        ("next_ibf", _u64p),
        ("user_bin", _u64p),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile, and, where the
    reference tree is present, oracle/_ref/robin_hood_ref (the reference's own
    robin_hood set driven by ref_robin_hood.cpp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.exists("/root/reference/include/hibf/contrib/robin_hood.hpp"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_hash_and_fit.restype = ctypes.c_uint64
        L.orc_hash_and_fit.argtypes = [ctypes.c_uint64] * 5
        L.orc_ibf_init.restype = ctypes.c_int
        L.orc_ibf_init.argtypes = [ctypes.POINTER(OrcIbf), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_ibf_emplace.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_uint64]
        L.orc_ibf_emplace_tracked.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
        L.orc_set_simd_counting.restype = ctypes.c_int
        L.orc_set_simd_counting.argtypes = [ctypes.c_int]
        L.orc_ibf_clear.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_uint64]
        L.orc_ibf_increase_bin_number_to.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_uint64]
        L.orc_ibf_increase_bin_number_to.restype = ctypes.c_int
        L.orc_ibf_bulk_contains.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_uint64,
                                            ctypes.c_void_p]
        L.orc_counting_vector_add.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_ibf_bulk_count.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_void_p, ctypes.c_uint64,
                                         ctypes.c_int, ctypes.c_void_p]
        L.orc_ibf_bulk_count_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_int]
        L.orc_ibf_membership_for.restype = ctypes.c_uint64
        L.orc_ibf_membership_for.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_void_p,
                                             ctypes.c_uint64, ctypes.c_uint16, ctypes.c_void_p]
        L.orc_hibf_membership_for.restype = ctypes.c_uint64
        L.orc_hibf_membership_for.argtypes = [ctypes.POINTER(OrcHibf), ctypes.c_void_p, ctypes.c_uint64,
                                              ctypes.c_uint16, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_hibf_membership_batch.restype = ctypes.c_uint64
        L.orc_hibf_membership_batch.argtypes = [ctypes.POINTER(OrcHibf), ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.c_void_p]
        L.orc_hibf_bulk_count.argtypes = [ctypes.POINTER(OrcHibf), ctypes.c_uint64, ctypes.c_void_p,
                                          ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        L.orc_hibf_bulk_count_batch.argtypes = [ctypes.POINTER(OrcHibf), ctypes.c_uint64, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_int]
        L.orc_hibf_membership_count_batch.argtypes = [ctypes.POINTER(OrcHibf), ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_int]
        L.orc_hll_wyhash.restype = ctypes.c_uint64
        L.orc_hll_wyhash.argtypes = [ctypes.c_uint64]
        L.orc_hll_add.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_hll_merge.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        L.orc_hll_estimate.restype = ctypes.c_double
        L.orc_hll_estimate.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.orc_hll_union_estimates.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                                              ctypes.c_void_p, ctypes.c_void_p]
        L.orc_mt19937_64.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_fs_new.restype = ctypes.c_void_p
        L.orc_fs_new.argtypes = []
        L.orc_fs_delete.argtypes = [ctypes.c_void_p]
        L.orc_fs_size.restype = ctypes.c_uint64
        L.orc_fs_size.argtypes = [ctypes.c_void_p]
        L.orc_fs_clear.argtypes = [ctypes.c_void_p]
        L.orc_fs_order.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_fs_emplace_synth.argtypes = [ctypes.c_void_p] + [ctypes.c_uint64] * 4
        L.orc_fs_emplace_values.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_fs_insert_set.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_fs_script.restype = ctypes.c_uint64
        L.orc_fs_script.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64]
        L.orc_synth_value.restype = ctypes.c_uint64
        L.orc_synth_value.argtypes = [ctypes.c_uint64] * 3
        L.orc_ibf_emplace_synthetic.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf)] + [ctypes.c_uint64] * 4 + \
            [ctypes.c_int]
        L.orc_ibf_emplace_synthetic_items.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_uint64] + \
            [ctypes.c_void_p] * 4 + [ctypes.c_uint64, ctypes.c_int]
        L.orc_ibf_emplace_parallel.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcIbf), ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
        L.orc_minimiser_hash.restype = ctypes.c_uint64
        L.orc_minimiser_hash.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                         ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64]
        _lib = L
    return _lib


_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def set_simd_counting(on: bool) -> bool:
    """Selects the AVX-512 restatement of counting_vector += for uint16_t
    counters (counting_vector.hpp:252-278) when the CPU has AVX-512BW;
    returns whether it is now in use."""
    return bool(lib().orc_set_simd_counting(1 if on else 0))


def hash_and_fit(h: int, seed: int, hash_shift: int, bin_size: int, technical_bins: int) -> int:
    return int(lib().orc_hash_and_fit(h, seed, hash_shift, bin_size, technical_bins))


def hash_and_fit_py(h: int, seed: int, hash_shift: int, bin_size: int, technical_bins: int) -> int:
    """Independent big-integer restatement of interleaved_bloom_filter.hpp:185-195."""
    m = (1 << 64) - 1
    h = (h * seed) & m
    h ^= h >> hash_shift
    h = (h * 11400714819323198485) & m
    h = (h * bin_size) >> 64
    return (h * technical_bins) & m


class Ibf:
    """An IBF held on the host: geometry plus its bit matrix as uint64 words."""

    def __init__(self, bins: int, bin_size: int, hash_funs: int = 2):
        self.g = OrcIbf()
        rc = lib().orc_ibf_init(ctypes.byref(self.g), bins, bin_size, hash_funs)
        if rc:
            raise ValueError(
                {1: "The number of bins must be > 0.",
                 2: "The number of hash functions must be > 0 and <= 5.",
                 3: "The size of a bin must be > 0."}[rc])
        self.words = np.zeros(self.g.word_count, dtype=np.uint64)

    @classmethod
    def from_words(cls, bins: int, bin_size: int, hash_funs: int, words: np.ndarray) -> "Ibf":
        ibf = cls.__new__(cls)
        ibf.g = geom(bins, bin_size, hash_funs)
        w = _u64(words)
        assert w.size == ibf.g.word_count
        ibf.words = w
        return ibf

    def emplace(self, values, bins) -> None:
        v = _u64(values)
        b = _u64(np.broadcast_to(np.asarray(bins, dtype=np.uint64), v.shape))
        lib().orc_ibf_emplace(_p(self.words), ctypes.byref(self.g), _p(v), _p(b), v.size)

    def emplace_tracked(self, values, bins, occupancy: np.ndarray) -> None:
        """emplace of the pairs in order with track_occupancy: occupancy (one
        uint64 per technical bin) grows by one for each value that set a new bit."""
        v = _u64(values)
        b = _u64(np.broadcast_to(np.asarray(bins, dtype=np.uint64), v.shape))
        assert occupancy.dtype == np.uint64 and occupancy.size >= self.g.technical_bins and occupancy.flags.c_contiguous
        lib().orc_ibf_emplace_tracked(_p(self.words), ctypes.byref(self.g), _p(v), _p(b), v.size, _p(occupancy))

    def emplace_parallel(self, values, bins, threads: int = 16) -> None:
        """emplace with `threads` threads (atomic OR; same result)."""
        v = _u64(values)
        b = _u64(np.broadcast_to(np.asarray(bins, dtype=np.uint64), v.shape))
        lib().orc_ibf_emplace_parallel(_p(self.words), ctypes.byref(self.g), _p(v), _p(b), v.size, threads)

    def emplace_synthetic(self, seed: int, first_bin: int, n_bins: int, per_bin: int, threads: int = 16) -> None:
        """Bin first_bin + b receives synth_value(seed, first_bin + b, j), j < per_bin."""
        lib().orc_ibf_emplace_synthetic(_p(self.words), ctypes.byref(self.g), seed, first_bin, n_bins, per_bin,
                                        threads)

    def emplace_synthetic_items(self, seed: int, sets, begins, counts, bins, threads: int = 16) -> None:
        """Item k: synth_value(seed, sets[k], begins[k] + j), j < counts[k], into bin bins[k]."""
        a = [_u64(x) for x in (sets, begins, counts, bins)]
        lib().orc_ibf_emplace_synthetic_items(_p(self.words), ctypes.byref(self.g), seed, *[_p(x) for x in a],
                                              a[0].size, threads)

    def clear(self, bin: int) -> None:
        lib().orc_ibf_clear(_p(self.words), ctypes.byref(self.g), bin)

    def increase_bin_number_to(self, new_bins: int) -> None:
        """interleaved_bloom_filter::increase_bin_number_to (re-strides in place)."""
        new_words = (new_bins + 63) // 64 * self.g.bin_size
        if new_words > self.words.size:
            self.words = np.concatenate([self.words, np.zeros(new_words - self.words.size, dtype=np.uint64)])
        if lib().orc_ibf_increase_bin_number_to(_p(self.words), ctypes.byref(self.g), new_bins):
            raise ValueError("The number of new bins must be >= the current number of bins.")

    def bulk_contains(self, value: int) -> np.ndarray:
        out = np.zeros(self.g.bin_words, dtype=np.uint64)
        lib().orc_ibf_bulk_contains(_p(self.words), ctypes.byref(self.g), int(value), _p(out))
        return out

    def contains_bits(self, value: int) -> list[int]:
        w = self.bulk_contains(value)
        return [int((int(w[b // 64]) >> (b % 64)) & 1) for b in range(self.g.bins)]

    def bulk_count(self, values, width: int = 2) -> np.ndarray:
        v = _u64(values)
        out = np.zeros(self.g.bins, dtype=_DT[width])
        lib().orc_ibf_bulk_count(_p(self.words), ctypes.byref(self.g), _p(v), v.size, width, _p(out))
        return out

    def bulk_count_batch(self, values, offsets, width: int = 2, threads: int = 1) -> np.ndarray:
        v, o = _u64(values), _u64(offsets)
        nq = o.size - 1
        out = np.zeros((nq, self.g.bins), dtype=_DT[width])
        lib().orc_ibf_bulk_count_batch(_p(self.words), ctypes.byref(self.g), _p(v), _p(o), nq, width, _p(out),
                                       threads)
        return out

    def membership_for(self, values, threshold: int) -> list[int]:
        v = _u64(values)
        out = np.zeros(max(self.g.bins, 1), dtype=np.uint64)
        n = lib().orc_ibf_membership_for(_p(self.words), ctypes.byref(self.g), _p(v), v.size, threshold, _p(out))
        return [int(x) for x in out[:n]]


def counting_vector_add(counts: np.ndarray, bits: np.ndarray, nbits: int) -> None:
    b = _u64(bits)
    lib().orc_counting_vector_add(_p(counts), counts.dtype.itemsize, _p(b), nbits)


class Hibf:
    """An HIBF as flat arrays: IBF geometries, one word arena and CSR tables."""

    def __init__(self, geoms, arena, word_offsets, tb_offsets, next_ibf, user_bin, n_user_bins):
        self.n_ibfs = len(geoms)
        self._geoms = (OrcIbf * max(self.n_ibfs, 1))()
        for i, g in enumerate(geoms):
            self._geoms[i] = g
        self.arena = _u64(arena)
        self.word_offsets = _u64(word_offsets)
        self.tb_offsets = _u64(tb_offsets)
        self.next_ibf = _u64(next_ibf)
        self.user_bin = _u64(user_bin)
        self.n_user_bins = int(n_user_bins)
        self.s = OrcHibf(self.n_ibfs, self._geoms,
                         self.arena.ctypes.data_as(_u64p), self.word_offsets.ctypes.data_as(_u64p),
                         self.tb_offsets.ctypes.data_as(_u64p), self.next_ibf.ctypes.data_as(_u64p),
                         self.user_bin.ctypes.data_as(_u64p))

    def membership_for(self, values, threshold: int) -> list[int]:
        v = _u64(values)
        cap = max(self.n_user_bins, 1)
        while True:
            # A user bin whose run is cut by deleted bins is reported once per run, so the list can be
            # longer than number_of_user_bins: the C side returns the full length and the call is repeated.
            out = np.zeros(cap, dtype=np.uint64)
            n = lib().orc_hibf_membership_for(ctypes.byref(self.s), _p(v), v.size, threshold, _p(out), cap)
            if n <= cap:
                return [int(x) for x in out[:n]]
            cap = int(n)

    def membership_batch(self, values, offsets, thresholds):
        v, o = _u64(values), _u64(offsets)
        nq = o.size - 1
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(thresholds, dtype=np.uint16), (nq,)))
        cap = 1 << 16
        while True:
            ids = np.zeros(cap, dtype=np.uint64)
            offs = np.zeros(nq + 1, dtype=np.uint64)
            total = lib().orc_hibf_membership_batch(ctypes.byref(self.s), _p(v), _p(o), nq, _p(t), _p(ids), cap,
                                                    _p(offs))
            if total <= cap:
                return ids[:total], offs
            cap = int(total)

    def bulk_count(self, values, threshold: int, width: int = 2) -> np.ndarray:
        v = _u64(values)
        out = np.zeros(self.n_user_bins, dtype=_DT[width])
        lib().orc_hibf_bulk_count(ctypes.byref(self.s), self.n_user_bins, _p(v), v.size, threshold, width, _p(out))
        return out

    def bulk_count_batch(self, values, offsets, thresholds, width: int = 2, threads: int = 1) -> np.ndarray:
        v, o = _u64(values), _u64(offsets)
        nq = o.size - 1
        t = _u64(np.broadcast_to(np.asarray(thresholds, dtype=np.uint64), (nq,)))
        out = np.zeros((nq, self.n_user_bins), dtype=_DT[width])
        lib().orc_hibf_bulk_count_batch(ctypes.byref(self.s), self.n_user_bins, _p(v), _p(o), nq, _p(t), width,
                                        _p(out), threads)
        return out

    def membership_hits_batch(self, values, offsets, thresholds, threads: int = 1) -> np.ndarray:
        v, o = _u64(values), _u64(offsets)
        nq = o.size - 1
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(thresholds, dtype=np.uint16), (nq,)))
        hits = np.zeros(nq, dtype=np.uint64)
        lib().orc_hibf_membership_count_batch(ctypes.byref(self.s), _p(v), _p(o), nq, _p(t), _p(hits), threads)
        return hits


class FlatSet:
    """The oracle's restatement of robin_hood::unordered_flat_set<uint64_t>, the
    set whose iteration order the reference chunks split user bins in."""

    def __init__(self):
        self.p = ctypes.c_void_p(lib().orc_fs_new())

    def __del__(self):
        try:
            lib().orc_fs_delete(self.p)
        except Exception:
            pass

    def __len__(self) -> int:
        return int(lib().orc_fs_size(self.p))

    def insert_synth(self, seed: int, set_: int, begin: int, count: int) -> None:
        lib().orc_fs_emplace_synth(self.p, seed, set_, begin, count)

    def insert_values(self, values) -> None:
        v = _u64(values)
        lib().orc_fs_emplace_values(self.p, _p(v), v.size)

    def insert_set(self, other: "FlatSet") -> None:
        lib().orc_fs_insert_set(self.p, other.p)

    def clear(self) -> None:
        lib().orc_fs_clear(self.p)

    def order(self) -> np.ndarray:
        out = np.zeros(max(len(self), 1), dtype=np.uint64)
        lib().orc_fs_order(self.p, _p(out))
        return out[: len(self)]


def flat_set_script(script: str) -> str:
    """Runs a golden-order script (the protocol of oracle/ref_robin_hood.cpp)
    on the oracle's restatement of robin_hood::unordered_flat_set<uint64_t>."""
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        need = lib().orc_fs_script(script.encode(), buf, cap)
        if need <= cap:
            return buf.value.decode()
        cap = int(need)


def synth_value(seed: int, bin: int, i: int) -> int:
    """Test-input generator (the synthetic stream of the device builds)."""
    return int(lib().orc_synth_value(seed, bin, i))


def geom(bins: int, bin_size: int, hash_funs: int) -> OrcIbf:
    g = OrcIbf()
    rc = lib().orc_ibf_init(ctypes.byref(g), bins, bin_size, hash_funs)
    if rc:
        raise ValueError(f"invalid IBF geometry (code {rc})")
    return g


_DNA4_RANK = np.zeros(256, dtype=np.uint64)
for _c, _r in (("C", 1), ("G", 2), ("T", 3), ("U", 3)):
    _DNA4_RANK[ord(_c)] = _r
    _DNA4_RANK[ord(_c.lower())] = _r


def kmer_hashes(seq, k: int) -> np.ndarray:
    """k-mer hashes of one DNA read as seqan3::views::kmer_hash(ungapped{k})
    over dna4 gives them to the reference's tutorial search
    (test/tutorial/search.cpp:37-41): for each window, the base-4 number of
    the dna4 ranks (A=0, C=1, G=2, T=U=3, anything else A), first letter most
    significant.  Restated from that definition (seqan3 is not available)."""
    b = np.frombuffer(seq.encode() if isinstance(seq, str) else bytes(seq), dtype=np.uint8)
    n = b.size - k + 1
    if n <= 0:
        return np.zeros(0, dtype=np.uint64)
    r = _DNA4_RANK[b]
    h = np.zeros(n, dtype=np.uint64)
    for j in range(k):
        h = (h << np.uint64(2)) | r[j:j + n]
    return h


# seqan3::views::minimiser_hash's default seed (seqan3 3.4.2, search/views/minimiser_hash.hpp).
MINIMISER_SEED = 0x8F3F73B5CF1C9ADE


def minimiser_hashes(seq, shape: int, window_size: int, seed: int = MINIMISER_SEED) -> np.ndarray:
    """Minimiser hashes of one DNA read as seqan3::views::minimiser_hash(shape,
    window_size{window_size}, seed{seed}) over dna4 computes them (restated in
    hibf_oracle.c, orc_minimiser_hash).  `shape` is seqan3's binary literal:
    the most significant set bit is the window's first position
    (ungapped{k} = 2^k - 1)."""
    if shape <= 0 or shape.bit_length() > 32:
        raise ValueError("shape must be non-zero with at most 32 positions")
    if window_size < shape.bit_length():
        raise ValueError("The size of the shape cannot be greater than the window size.")
    b = np.frombuffer(seq.encode() if isinstance(seq, str) else bytes(seq), dtype=np.uint8)
    n = max(0, b.size - shape.bit_length() + 1)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    got = lib().orc_minimiser_hash(_p(b) if b.size else None, b.size, shape, window_size, seed, _p(out), out.size)
    assert got <= out.size
    return out[:got].copy()


def kmer_threshold(n_kmers: int, fraction: float) -> int:
    """uint16_t(fraction * number of k-mers) (test/tutorial/search.cpp:42)."""
    t = fraction * n_kmers
    return 65535 if t >= 65535 else max(0, int(t))


# ------------------------------------------------- HyperLogLog (build side) --
# src/sketch/hyperloglog.cpp restated in hibf_oracle.c (orc_hll_*).

def hll_wyhash(value: int) -> int:
    return int(lib().orc_hll_wyhash(value))


def hll_sketch(values, bits: int = 5) -> np.ndarray:
    """Registers of hyperloglog(bits) after add() of every value."""
    regs = np.zeros(1 << bits, dtype=np.uint8)
    v = _u64(values)
    lib().orc_hll_add(_p(regs), bits, _p(v), v.size)
    return regs


def hll_merge(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.uint8).copy()
    other = np.ascontiguousarray(b, dtype=np.uint8)
    lib().orc_hll_merge(_p(out), _p(other), int(out.size).bit_length() - 1)
    return out


def hll_estimate(regs: np.ndarray) -> float:
    r = np.ascontiguousarray(regs, dtype=np.uint8)
    return float(lib().orc_hll_estimate(_p(r), int(r.size).bit_length() - 1))


def hll_union_estimates(registers: np.ndarray, positions) -> np.ndarray:
    """Union estimates of the layout DP (toolbox.cpp:35-51) for every column j
    and k < j of `positions` over sketches `registers` (n x 2^bits uint8),
    packed by column: column j's j values start at j (j - 1) / 2."""
    regs = np.ascontiguousarray(registers, dtype=np.uint8)
    bits = int(regs.shape[1]).bit_length() - 1
    pos = _u64(positions)
    np_ = pos.size
    out = np.zeros(max(np_ * (np_ - 1) // 2, 1), dtype=np.uint64)
    scratch = np.zeros(regs.shape[1], dtype=np.uint8)
    if np_ >= 2:
        lib().orc_hll_union_estimates(_p(regs), bits, _p(pos), np_, _p(scratch), _p(out))
    return out[:np_ * (np_ - 1) // 2]


def mt19937_64(seed: int, n: int) -> np.ndarray:
    """The first n outputs of std::mt19937_64(seed)."""
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_mt19937_64(seed, _p(out), n)
    return out
