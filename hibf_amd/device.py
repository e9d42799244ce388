"""Device-resident IBF / HIBF handles over the C-ABI, with torch tensors as HBM buffers.

Mirrors the reference's query interface (names and argument meaning of
seqan::hibf::interleaved_bloom_filter and hierarchical_interleaved_bloom_filter,
interleaved_bloom_filter.hpp:225-446 and hierarchical_interleaved_bloom_filter.hpp:197-247),
batched: a batch of queries is a CSR pair (values, offsets) of uint64 tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import native
from .native import Geometry, check, lib

_DTYPES = {1: torch.uint8, 2: torch.uint16, 4: torch.uint32, 8: torch.uint64}


def _release(obj, destroy: str) -> None:
    """Frees a handle from __del__; at interpreter exit the binding may already
    be torn down, and the process's end releases the device memory anyway."""
    h = getattr(obj, "h", None)
    if not h:
        return
    obj.h = None
    try:
        getattr(lib(), destroy)(h)
    except Exception:
        pass


def _dev() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("hibf_amd needs a HIP device; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def as_device_u64(x) -> torch.Tensor:
    """uint64 tensor on the device from a numpy array, list or tensor."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.dtype != torch.uint64:
            t = t.to(torch.int64).view(torch.uint64) if t.dtype != torch.int64 else t.view(torch.uint64)
        return t.to(_dev()).contiguous()
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64))
    if not a.flags.writeable:
        a = a.copy()
    return torch.from_numpy(a.view(np.int64)).to(_dev()).view(torch.uint64)


def _check_buffer(t, name: str, min_numel: int, element_size: int | None = 8, exact: bool = False) -> None:
    """Raises ValueError unless t is a contiguous device tensor of at least
    (exactly, with exact=True) min_numel elements of element_size bytes: the
    C-ABI reads and writes these buffers by count, without their sizes."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous device tensor")
    if element_size is not None and t.element_size() != element_size:
        raise ValueError(f"{name} must have {element_size}-byte elements (got {t.element_size()})")
    if (t.numel() != min_numel) if exact else (t.numel() < min_numel):
        raise ValueError(f"{name} must have {'exactly' if exact else 'at least'} {min_numel} elements "
                         f"(got {t.numel()})")


def to_numpy_u64(t: torch.Tensor) -> np.ndarray:
    return t.view(torch.int64).cpu().numpy().view(np.uint64)


def reads_csr(reads) -> tuple[np.ndarray, np.ndarray]:
    """(bytes, offsets) of a list of DNA reads given as str or bytes."""
    parts = [r.encode() if isinstance(r, str) else bytes(r) for r in reads]
    offsets = np.zeros(len(parts) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum([len(p) for p in parts], dtype=np.uint64)
    data = np.frombuffer(b"".join(parts), dtype=np.uint8) if parts and offsets[-1] else np.zeros(0, dtype=np.uint8)
    return data, offsets


def kmer_hash(reads, k: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Device k-mer hashes of DNA reads (hibf_gpu_kmer_hash): (hashes, offsets)
    as uint64 device tensors in CSR form, one row of hashes per read.  `reads`
    is a list of str/bytes or a (bases uint8, offsets uint64) CSR pair."""
    data, offsets = reads if isinstance(reads, tuple) else reads_csr(reads)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    dev = _dev()
    d_seq = torch.from_numpy(np.ascontiguousarray(data).copy()).to(dev) if data.size else \
        torch.zeros(1, dtype=torch.uint8, device=dev)
    d_off = as_device_u64(offsets)
    n = offsets.size - 1
    out_off = torch.empty(n + 1, dtype=torch.uint64, device=dev)
    total = ctypes.c_uint64()
    cap = max(int(offsets[-1]), 1)
    hashes = torch.empty(cap, dtype=torch.uint64, device=dev)
    check(lib().hibf_gpu_kmer_hash(_ptr(d_seq), _ptr(d_off), n, k, _ptr(out_off), _ptr(hashes), cap,
                                   ctypes.byref(total), _stream()))
    return hashes[: total.value], out_off


# seqan3::views::minimiser_hash's default seed.
MINIMISER_SEED = 0x8F3F73B5CF1C9ADE


def minimiser_hash(reads, shape: int, window_size: int, seed: int = MINIMISER_SEED) -> tuple[torch.Tensor, torch.Tensor]:
    """Device minimiser hashes of DNA reads (hibf_gpu_minimiser_hash), as
    seqan3::views::minimiser_hash(shape, window_size{window_size}, seed{seed})
    computes them: (hashes, offsets) as uint64 device tensors in CSR form.
    `shape` is seqan3's binary literal (most significant set bit = first
    position; ungapped{k} = 2^k - 1).  `reads` is a list of str/bytes or a
    (bases uint8, offsets uint64) CSR pair."""
    data, offsets = reads if isinstance(reads, tuple) else reads_csr(reads)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    dev = _dev()
    d_seq = torch.from_numpy(np.ascontiguousarray(data).copy()).to(dev) if data.size else \
        torch.zeros(1, dtype=torch.uint8, device=dev)
    d_off = as_device_u64(offsets)
    n = offsets.size - 1
    out_off = torch.empty(n + 1, dtype=torch.uint64, device=dev)
    total = ctypes.c_uint64()
    # At most one minimiser per k-mer position: the base count bounds them.
    cap = max(int(offsets[-1]), 1)
    hashes = torch.empty(cap, dtype=torch.uint64, device=dev)
    check(lib().hibf_gpu_minimiser_hash(_ptr(d_seq), _ptr(d_off), n, ctypes.c_uint64(shape), window_size,
                                        ctypes.c_uint64(seed), _ptr(out_off), _ptr(hashes), cap,
                                        ctypes.byref(total), _stream()))
    return hashes[: total.value], out_off


def hll_sketches(sets, bits: int = 12) -> torch.Tensor:
    """HyperLogLog sketches of value sets on the device (hibf_gpu_hll_sketch):
    a (len(sets), 2^bits) uint8 device tensor, one row of registers per set
    (seqan::hibf::sketch::hyperloglog, one per user bin as compute_sketches
    builds them).  `sets` is a list of uint64 sequences or a (values, offsets)
    CSR pair."""
    values, offsets = sets if isinstance(sets, tuple) else csr(sets)
    n = int(np.asarray(offsets).size - 1) if not torch.is_tensor(offsets) else offsets.numel() - 1
    v, o = as_device_u64(values), as_device_u64(offsets)
    regs = torch.empty((max(n, 1), 1 << bits), dtype=torch.uint8, device=o.device)
    check(lib().hibf_gpu_hll_sketch(_ptr(v) if v.numel() else None, _ptr(o), n, bits, _ptr(regs), _stream()))
    return regs[:n]


def hll_sketches_synthetic(sizes, seed: int, bits: int = 12) -> torch.Tensor:
    """Sketches of the synthetic user bins of DeviceHibf.build_synthetic(sizes, seed)."""
    s = np.ascontiguousarray(np.asarray(sizes, dtype=np.uint64))
    regs = torch.empty((max(s.size, 1), 1 << bits), dtype=torch.uint8, device=_dev())
    check(lib().hibf_gpu_hll_sketch_synthetic(int(seed), s.ctypes.data_as(ctypes.c_void_p), s.size, bits, _ptr(regs),
                                              _stream()))
    return regs[: s.size]


def hll_merge(registers: torch.Tensor, groups) -> torch.Tensor:
    """Register-wise max of each group of sketches (hibf_gpu_hll_merge);
    `groups` is a list of lists of sketch rows."""
    bits = int(registers.shape[1]).bit_length() - 1
    members, offsets = csr([np.asarray(g, dtype=np.uint64) for g in groups])
    out = torch.empty((max(len(groups), 1), registers.shape[1]), dtype=torch.uint8, device=registers.device)
    m, o = as_device_u64(members if members.size else np.zeros(1, dtype=np.uint64)), as_device_u64(offsets)
    check(lib().hibf_gpu_hll_merge(_ptr(registers), bits, _ptr(m), _ptr(o), len(groups), _ptr(out), _stream()))
    return out[: len(groups)]


def hll_estimates(registers: torch.Tensor) -> np.ndarray:
    """hyperloglog::estimate of every row of a register tensor (float64)."""
    bits = int(registers.shape[1]).bit_length() - 1
    n = registers.shape[0]
    out = np.zeros(max(n, 1), dtype=np.float64)
    regs = registers.contiguous()
    check(lib().hibf_gpu_hll_estimate(_ptr(regs) if n else None, n, bits, out.ctypes.data_as(ctypes.c_void_p),
                                      _stream()))
    return out[:n]


def hll_union_estimates(registers: torch.Tensor, positions) -> torch.Tensor:
    """Union estimates of the layout DP (hibf_gpu_hll_union_estimates) for the
    sketches `registers` (n x 2^bits uint8, device) taken in `positions` order:
    a uint64 device tensor, column j's j values (k = 0 .. j - 1) starting at
    j (j - 1) / 2."""
    bits = int(registers.shape[1]).bit_length() - 1
    pos = as_device_u64(np.asarray(positions, dtype=np.uint64))
    npos = pos.numel()
    out = torch.empty(max(npos * (npos - 1) // 2, 1), dtype=torch.uint64, device=registers.device)
    regs = registers.contiguous()
    check(lib().hibf_gpu_hll_union_estimates(_ptr(regs), regs.shape[0], bits, _ptr(pos), npos, _ptr(out), _stream()))
    return out[: npos * (npos - 1) // 2]


def csr(queries) -> tuple[np.ndarray, np.ndarray]:
    """(values, offsets) for a list of uint64 sequences."""
    lens = [len(q) for q in queries]
    offsets = np.zeros(len(queries) + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(lens, dtype=np.uint64)
    values = np.concatenate([np.asarray(q, dtype=np.uint64) for q in queries]) if queries and sum(lens) else \
        np.zeros(0, dtype=np.uint64)
    return values, offsets


class Agent:
    """Per-caller device workspace (the reference's one-agent-per-thread rule)."""

    def __init__(self, max_queries_per_launch: int = 0):
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_agent_create(max_queries_per_launch, ctypes.byref(h)))
        self.h = h

    def stats(self, reset: bool = False) -> dict:
        """IBF visits, lookups and row bytes of HIBF descents since the last reset."""
        v, lk, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().hibf_gpu_agent_stats(self.h, ctypes.byref(v), ctypes.byref(lk), ctypes.byref(b), int(reset)))
        return {"visits": int(v.value), "lookups": int(lk.value), "row_bytes": int(b.value)}

    def status(self) -> tuple[int, bool]:
        """(ids found by the last async membership launch, whether any launch dropped ids)."""
        t, d = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().hibf_gpu_agent_status(self.h, ctypes.byref(t), ctypes.byref(d)))
        return int(t.value), bool(d.value)

    def __del__(self):
        _release(self, "hibf_gpu_agent_destroy")


class DeviceIbf:
    """interleaved_bloom_filter(bin_count, bin_size, hash_function_count) in HBM."""

    def __init__(self, bin_count: int, bin_size: int, hash_function_count: int = 2):
        _dev()
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_ibf_create(bin_count, bin_size, hash_function_count, ctypes.byref(h)))
        self.h = h
        g = Geometry()
        check(lib().hibf_gpu_ibf_get_geometry(self.h, ctypes.byref(g)))
        self.geometry = g
        self._agent = None

    def __del__(self):
        _release(self, "hibf_gpu_ibf_destroy")

    @classmethod
    def load(cls, path: str, offset: int = 0) -> tuple["DeviceIbf", int]:
        """Loads an interleaved_bloom_filter archive (cereal binary layout)
        starting at byte `offset`; returns the IBF and the offset after it."""
        _dev()
        h, end = ctypes.c_void_p(), ctypes.c_uint64()
        check(lib().hibf_gpu_ibf_load(str(path).encode(), offset, ctypes.byref(h), ctypes.byref(end)))
        self = cls.__new__(cls)
        self.h = h
        g = Geometry()
        check(lib().hibf_gpu_ibf_get_geometry(self.h, ctypes.byref(g)))
        self.geometry = g
        self._agent = None
        return self, int(end.value)

    def save(self, path: str, append: bool = False) -> None:
        """Writes the interleaved_bloom_filter archive (cereal binary layout)."""
        check(lib().hibf_gpu_ibf_save(self.h, str(path).encode(), int(append)))

    def try_increase_bin_number_to(self, new_bins: int) -> bool:
        ok = ctypes.c_int()
        check(lib().hibf_gpu_ibf_try_increase_bin_number_to(self.h, new_bins, ctypes.byref(ok)))
        check(lib().hibf_gpu_ibf_get_geometry(self.h, ctypes.byref(self.geometry)))
        return bool(ok.value)

    def increase_bin_number_to(self, new_bins: int) -> None:
        """interleaved_bloom_filter::increase_bin_number_to, re-strided on the device."""
        check(lib().hibf_gpu_ibf_increase_bin_number_to(self.h, new_bins))
        check(lib().hibf_gpu_ibf_get_geometry(self.h, ctypes.byref(self.geometry)))

    def occupancy(self) -> tuple[np.ndarray, bool]:
        occ = np.zeros(int(self.geometry.technical_bins), dtype=np.uint64)
        track = ctypes.c_int()
        check(lib().hibf_gpu_ibf_occupancy(self.h, occ.ctypes.data_as(ctypes.c_void_p), ctypes.byref(track)))
        return occ, bool(track.value)

    def set_occupancy(self, occupancy, track: bool) -> None:
        occ = np.ascontiguousarray(occupancy, dtype=np.uint64)
        if occ.size != int(self.geometry.technical_bins):
            raise ValueError("occupancy needs technical_bins entries")
        check(lib().hibf_gpu_ibf_set_occupancy(self.h, occ.ctypes.data_as(ctypes.c_void_p), int(track)))

    # interleaved_bloom_filter.hpp:400-427
    def bin_count(self) -> int:
        return int(self.geometry.bins)

    def bin_size(self) -> int:
        return int(self.geometry.bin_size)

    def hash_function_count(self) -> int:
        return int(self.geometry.hash_funs)

    def bit_size(self) -> int:
        return int(self.geometry.technical_bins * self.geometry.bin_size)

    def emplace(self, values, bins) -> None:
        v = as_device_u64(values)
        if isinstance(bins, (int, np.integer)):
            check(lib().hibf_gpu_ibf_emplace(self.h, _ptr(v), None, int(bins), v.numel(), _stream()))
        else:
            b = as_device_u64(bins)
            assert b.numel() == v.numel()
            check(lib().hibf_gpu_ibf_emplace(self.h, _ptr(v), _ptr(b), 0, v.numel(), _stream()))
        self._keep = (v,)
        torch.cuda.current_stream().synchronize()

    def emplace_synthetic(self, seed: int, first_bin: int, n_bins: int, values_per_bin: int) -> None:
        """Technical bin first_bin+b receives synth_values(seed, first_bin+b, 0..values_per_bin)."""
        check(lib().hibf_gpu_ibf_emplace_synthetic(self.h, seed, first_bin, n_bins, values_per_bin, _stream()))
        torch.cuda.current_stream().synchronize()

    def device_words(self) -> int:
        p = ctypes.c_void_p()
        check(lib().hibf_gpu_ibf_device_words(self.h, ctypes.byref(p)))
        return int(p.value)

    def bulk_count_into(self, values: torch.Tensor, offsets: torch.Tensor, out: torch.Tensor,
                        counter_bytes: int = 2, stream=None) -> None:
        """bulk_count into a preallocated (nq, bin_count) tensor, asynchronously.
        values must hold the offsets[nq] values the offsets address."""
        nq = offsets.numel() - 1
        _check_buffer(offsets, "offsets", 1)
        _check_buffer(values, "values", 1 if nq else 0)
        if counter_bytes not in _DTYPES:
            raise ValueError("counter_bytes must be 1, 2, 4 or 8")
        _check_buffer(out, "out", nq * self.bin_count(), counter_bytes)
        s = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
        check(lib().hibf_gpu_ibf_bulk_count(self.h, _ptr(values), _ptr(offsets), nq, counter_bytes, _ptr(out), s))

    def clear(self, bins) -> None:
        b = np.ascontiguousarray(np.atleast_1d(np.asarray(bins, dtype=np.uint64)))
        check(lib().hibf_gpu_ibf_clear(self.h, b.ctypes.data_as(ctypes.c_void_p), b.size, _stream()))

    def download(self) -> np.ndarray:
        out = np.zeros(self.geometry.n_words, dtype=np.uint64)
        check(lib().hibf_gpu_ibf_download(self.h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def upload(self, words: np.ndarray) -> None:
        w = np.ascontiguousarray(words, dtype=np.uint64)
        check(lib().hibf_gpu_ibf_upload(self.h, w.ctypes.data_as(ctypes.c_void_p), w.size))

    def bulk_contains(self, values) -> torch.Tensor:
        v = as_device_u64(values)
        out = torch.empty((v.numel(), int(self.geometry.bin_words)), dtype=torch.uint64, device=v.device)
        check(lib().hibf_gpu_ibf_bulk_contains(self.h, _ptr(v), v.numel(), _ptr(out), _stream()))
        return out

    def bulk_contains_into(self, values: torch.Tensor, out: torch.Tensor) -> None:
        """bulk_contains into a preallocated (n, bin_words) uint64 device tensor."""
        n = values.numel()
        if values.dtype not in (torch.uint64, torch.int64) or not values.is_contiguous():
            raise ValueError("values must be a contiguous 64-bit device tensor")
        if out.dtype not in (torch.uint64, torch.int64) or out.numel() < n * int(self.geometry.bin_words) \
                or not out.is_contiguous():
            raise ValueError("out must be a contiguous 64-bit tensor of at least n * bin_words words")
        check(lib().hibf_gpu_ibf_bulk_contains(self.h, _ptr(values), n, _ptr(out), _stream()))

    def bulk_count(self, values, offsets, counter_bytes: int = 2) -> torch.Tensor:
        v = as_device_u64(values)
        o = as_device_u64(offsets)
        nq = o.numel() - 1
        out = torch.empty((nq, self.bin_count()), dtype=_DTYPES[counter_bytes], device=o.device)
        check(lib().hibf_gpu_ibf_bulk_count(self.h, _ptr(v), _ptr(o), nq, counter_bytes, _ptr(out), _stream()))
        return out

    def membership_for(self, values, offsets, thresholds, agent: Agent | None = None):
        """Returns (ids, hit_offsets) as device tensors (CSR over queries)."""
        v = as_device_u64(values)
        o = as_device_u64(offsets)
        nq = o.numel() - 1
        t = _thresholds16(thresholds, nq, o.device)
        agent = agent or self._default_agent()
        return _membership(lambda ids, cap, hoff, total: lib().hibf_gpu_ibf_membership_for(
            self.h, agent.h, _ptr(v), _ptr(o), nq, _ptr(t), _ptr(ids), cap, _ptr(hoff), ctypes.byref(total),
            _stream()), nq, o.device)

    def _default_agent(self) -> Agent:
        if self._agent is None:
            self._agent = Agent()
        return self._agent


def _thresholds16(thresholds, nq: int, device) -> torch.Tensor:
    if isinstance(thresholds, torch.Tensor) and thresholds.is_cuda:
        if thresholds.dtype != torch.uint16 or thresholds.numel() != nq or not thresholds.is_contiguous():
            raise ValueError("device thresholds must be a contiguous uint16 tensor of one entry per query")
        return thresholds
    a = np.array(np.broadcast_to(np.asarray(thresholds, dtype=np.uint16), (nq,)), dtype=np.uint16)  # own copy
    return torch.from_numpy(a.view(np.int16)).to(device).view(torch.uint16)


def _membership(call, nq: int, device):
    cap = max(1024, nq * 4)
    for _ in range(2):
        ids = torch.empty(cap, dtype=torch.uint64, device=device)
        hoff = torch.empty(nq + 1, dtype=torch.uint64, device=device)
        total = ctypes.c_uint64(0)
        rc = call(ids, cap, hoff, total)
        if rc == native.ERR_CAPACITY:
            cap = int(total.value)
            continue
        check(rc)
        torch.cuda.current_stream().synchronize()
        return ids[: int(total.value)], hoff
    raise native.HibfGpuError(native.ERR_CAPACITY, "membership output did not fit after resizing")


class DeviceHibf:
    """A built HIBF uploaded to HBM: IBF arena + routing tables
    (hierarchical_interleaved_bloom_filter.hpp:145, 209-238)."""

    def __init__(self, geoms, word_offsets, tb_offsets, next_ibf, user_bin, number_of_user_bins: int,
                 arena: np.ndarray | None = None):
        _dev()
        n = len(geoms)
        garr = (Geometry * n)()
        for i, (bins, bin_size, h) in enumerate(geoms):
            garr[i].bins, garr[i].bin_size, garr[i].hash_funs = bins, bin_size, h
        arrs = [np.ascontiguousarray(np.asarray(x, dtype=np.uint64)) for x in
                (word_offsets, tb_offsets, next_ibf, user_bin)]
        ar = None if arena is None else np.ascontiguousarray(arena, dtype=np.uint64)
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_hibf_create(n, garr, None if ar is None else ar.ctypes.data_as(ctypes.c_void_p),
                                         *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs],
                                         number_of_user_bins, ctypes.byref(h)))
        self.h = h
        self.number_of_user_bins = int(number_of_user_bins)
        self._agent = None

    @classmethod
    def _from_handle(cls, h: ctypes.c_void_p) -> "DeviceHibf":
        self = cls.__new__(cls)
        self.h = h
        n_ub = ctypes.c_uint64()
        check(lib().hibf_gpu_hibf_info(self.h, None, None, None, ctypes.byref(n_ub), None))
        self.number_of_user_bins = int(n_ub.value)
        self._agent = None
        return self

    @classmethod
    def load(cls, path: str, offset: int = 0) -> tuple["DeviceHibf", int]:
        """Loads a hierarchical_interleaved_bloom_filter archive (cereal binary
        layout) starting at byte `offset`; returns the HIBF and the offset after it."""
        _dev()
        h, end = ctypes.c_void_p(), ctypes.c_uint64()
        check(lib().hibf_gpu_hibf_load(str(path).encode(), offset, ctypes.byref(h), ctypes.byref(end)))
        return cls._from_handle(h), int(end.value)

    def save(self, path: str, append: bool = False) -> None:
        """Writes the hierarchical_interleaved_bloom_filter archive (cereal binary layout)."""
        check(lib().hibf_gpu_hibf_save(self.h, str(path).encode(), int(append)))

    @classmethod
    def build(cls, user_bins, layout_text: str | None = None, **config) -> "DeviceHibf":
        """hierarchical_interleaved_bloom_filter(config) over explicit user bins
        (a list of uint64 sequences); config keywords as seqan::hibf::config."""
        _dev()
        values, offsets = csr(user_bins)
        c = native.make_config(number_of_user_bins=len(user_bins), **config)
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_hibf_build(ctypes.byref(c), offsets.ctypes.data_as(ctypes.c_void_p),
                                        values.ctypes.data_as(ctypes.c_void_p),
                                        None if layout_text is None else layout_text.encode(), ctypes.byref(h)))
        return cls._from_handle(h)

    @classmethod
    def build_synthetic(cls, sizes, seed: int, layout_text: str | None = None, **config) -> "DeviceHibf":
        """HIBF over synthetic user bins: user bin u holds synth_values(seed, u, 0..sizes[u])."""
        _dev()
        sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.uint64))
        c = native.make_config(number_of_user_bins=sz.size, **config)
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_hibf_build_synthetic(ctypes.byref(c), sz.ctypes.data_as(ctypes.c_void_p), seed,
                                                  None if layout_text is None else layout_text.encode(),
                                                  ctypes.byref(h)))
        return cls._from_handle(h)

    def tables(self) -> dict:
        """Host copy of the routing tables and IBF geometries."""
        n_ibfs, n_tbs, n_words, n_ub, depth = (ctypes.c_uint64() for _ in range(5))
        check(lib().hibf_gpu_hibf_info(self.h, ctypes.byref(n_ibfs), ctypes.byref(n_tbs), ctypes.byref(n_words),
                                       ctypes.byref(n_ub), ctypes.byref(depth)))
        n, t = n_ibfs.value, n_tbs.value
        geoms = (Geometry * n)()
        wo = np.zeros(n, dtype=np.uint64)
        to = np.zeros(n + 1, dtype=np.uint64)
        nx = np.zeros(t, dtype=np.uint64)
        ub = np.zeros(t, dtype=np.uint64)
        check(lib().hibf_gpu_hibf_export_tables(self.h, geoms, *[a.ctypes.data_as(ctypes.c_void_p)
                                                                 for a in (wo, to, nx, ub)]))
        return {"geoms": [(g.bins, g.bin_size, g.hash_funs) for g in geoms], "word_offsets": wo, "tb_offsets": to,
                "next_ibf": nx, "user_bin": ub, "n_words": n_words.value, "n_user_bins": n_ub.value,
                "depth": depth.value}

    def layout_text(self) -> str:
        """The layout this HIBF was built from (reference text format; empty
        when it was created from tables or loaded)."""
        need = ctypes.c_uint64()
        rc = lib().hibf_gpu_hibf_layout(self.h, None, 0, ctypes.byref(need))
        if rc not in (native.OK, native.ERR_CAPACITY):
            check(rc)
        buf = ctypes.create_string_buffer(int(need.value))
        check(lib().hibf_gpu_hibf_layout(self.h, buf, need.value, ctypes.byref(need)))
        return buf.value.decode()

    def __del__(self):
        _release(self, "hibf_gpu_hibf_destroy")

    def emplace(self, ibf_idx: int, values, bins) -> None:
        v = as_device_u64(values)
        if isinstance(bins, (int, np.integer)):
            check(lib().hibf_gpu_hibf_emplace(self.h, ibf_idx, _ptr(v), None, int(bins), v.numel(), _stream()))
        else:
            b = as_device_u64(bins)
            check(lib().hibf_gpu_hibf_emplace(self.h, ibf_idx, _ptr(v), _ptr(b), 0, v.numel(), _stream()))
        torch.cuda.current_stream().synchronize()

    def ibf(self, i: int) -> DeviceIbf:
        """ibf_vector[i]: a DeviceIbf over IBF i's words in the arena (no copy;
        it keeps this HIBF alive)."""
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_hibf_ibf_view(self.h, int(i), ctypes.byref(h)))
        view = DeviceIbf.__new__(DeviceIbf)
        view.h = h
        g = Geometry()
        check(lib().hibf_gpu_ibf_get_geometry(view.h, ctypes.byref(g)))
        view.geometry = g
        view._agent = None
        view._owner = self
        return view

    def download_arena(self) -> np.ndarray:
        n = ctypes.c_uint64()
        check(lib().hibf_gpu_hibf_arena_words(self.h, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint64)
        check(lib().hibf_gpu_hibf_download_arena(self.h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def membership_for(self, values, offsets, thresholds, agent: Agent | None = None):
        v = as_device_u64(values)
        o = as_device_u64(offsets)
        nq = o.numel() - 1
        t = _thresholds16(thresholds, nq, o.device)
        agent = agent or self._default_agent()
        return _membership(lambda ids, cap, hoff, total: lib().hibf_gpu_hibf_membership_for(
            self.h, agent.h, _ptr(v), _ptr(o), nq, _ptr(t), _ptr(ids), cap, _ptr(hoff), ctypes.byref(total),
            _stream()), nq, o.device)

    def membership_for_reads(self, reads, k: int, fraction: float = 0.7, agent: Agent | None = None,
                             minimiser: tuple | None = None):
        """The tutorial search (test/tutorial/search.cpp:37-45) on the device:
        each read's k-mer hashes are queried with threshold
        uint16_t(fraction * number of k-mers).  With minimiser = (shape,
        window_size) or (shape, window_size, seed), the read's minimiser hashes
        (minimiser_hash) are queried instead, with threshold
        uint16_t(fraction * number of minimisers), and k is not used.  Returns
        one id list per read, in the reference's depth-first order."""
        hashes, offsets = kmer_hash(reads, k) if minimiser is None else minimiser_hash(reads, *minimiser)
        n = offsets.numel() - 1
        thr = torch.empty(max(n, 1), dtype=torch.uint16, device=offsets.device)
        check(lib().hibf_gpu_kmer_thresholds(_ptr(offsets), n, float(fraction), _ptr(thr), _stream()))
        if hashes.numel() == 0:
            hashes = torch.zeros(1, dtype=torch.uint64, device=offsets.device)
        ids, hoff = self.membership_for(hashes, offsets, thr[:n], agent)
        ids, hoff = to_numpy(ids), to_numpy(hoff)
        return [ids[int(hoff[i]):int(hoff[i + 1])].tolist() for i in range(n)]

    def membership_prepare(self, agent: Agent, max_queries: int) -> None:
        check(lib().hibf_gpu_hibf_membership_prepare(self.h, agent.h, max_queries))

    def membership_launch(self, agent: Agent, values: torch.Tensor, offsets: torch.Tensor,
                          thresholds: torch.Tensor, ids: torch.Tensor, hit_start: torch.Tensor,
                          hit_count: torch.Tensor, stream=None) -> None:
        """Asynchronous membership over the CSR batch (offsets has nq+1 entries);
        query q's ids are ids[hit_start[q] : hit_start[q] + hit_count[q]]."""
        nq = offsets.numel() - 1
        _check_buffer(offsets, "offsets", 1)
        _check_buffer(values, "values", 1 if nq else 0)
        _check_buffer(thresholds, "thresholds", nq, 2, exact=True)
        _check_buffer(ids, "ids", 0)
        _check_buffer(hit_start, "hit_start", nq)
        _check_buffer(hit_count, "hit_count", nq)
        s = _stream() if stream is None else ctypes.c_void_p(stream.cuda_stream)
        check(lib().hibf_gpu_hibf_membership_launch(self.h, agent.h, _ptr(values), _ptr(offsets), nq,
                                                    _ptr(thresholds), _ptr(ids), ids.numel(), _ptr(hit_start),
                                                    _ptr(hit_count), s))

    def serve(self, values: np.ndarray, offsets: np.ndarray, thresholds: np.ndarray, arrival_s: np.ndarray,
              window_us: float = 500.0, max_batch: int = 65536, max_batch_values: int = 1 << 24,
              budget_bytes: int = 4 << 30, streams: int = 4, results_per_query: int = 16):
        """Runs the serving loop over requests arriving at arrival_s (host arrays).
        values may be a numpy array or a pinned CPU tensor (torch pin_memory),
        which the loop copies to the device without staging.
        Returns (latency_s, ids[nq, R], counts, stats dict)."""
        if isinstance(values, torch.Tensor):
            if values.is_cuda or values.dtype not in (torch.int64, torch.uint64) or not values.is_contiguous():
                raise ValueError("values must be a contiguous 64-bit host tensor")
            v = values
        else:
            v = np.ascontiguousarray(values, dtype=np.uint64)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        t = np.ascontiguousarray(thresholds, dtype=np.uint16)
        a = np.ascontiguousarray(arrival_s, dtype=np.float64)
        nq = o.size - 1
        if nq < 0 or t.size != nq or a.size != nq:
            raise ValueError("offsets need nq + 1 entries and thresholds and arrival_s one per request")
        if (v.numel() if isinstance(v, torch.Tensor) else v.size) < (int(o[-1]) if nq >= 0 else 0):
            raise ValueError("values hold fewer elements than offsets[-1] addresses")
        lat = np.zeros(nq, dtype=np.float64)
        ids = np.zeros((nq, results_per_query), dtype=np.uint64)
        cnt = np.zeros(nq, dtype=np.uint32)
        p = native.ServeParams(window_us, max_batch, max_batch_values, budget_bytes, streams, results_per_query)
        st = native.ServeStats()
        P = lambda x: ctypes.c_void_p(x.data_ptr()) if isinstance(x, torch.Tensor) \
            else x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().hibf_gpu_hibf_serve(self.h, ctypes.byref(p), P(v), P(o), P(t), P(a), nq, P(lat), P(ids), P(cnt),
                                        ctypes.byref(st)))
        stats = {k: getattr(st, k) for k, _ in native.ServeStats._fields_}
        return lat, ids, cnt, stats

    def serve_reads(self, reads, arrival_s: np.ndarray, k: int = 19, fraction: float = 0.7, window_us: float = 500.0,
                    max_batch: int = 65536, max_batch_bases: int = 1 << 26, budget_bytes: int = 4 << 30,
                    streams: int = 4, results_per_query: int = 16, pinned: bool = True,
                    minimiser: tuple | None = None):
        """The serving loop for reads (hibf_gpu_hibf_serve_reads): reads (str or
        bytes, or a (bases, offsets) pair) arrive at arrival_s; each is hashed
        on the device and answered with threshold uint16_t(fraction * #k-mers).
        With minimiser = (shape, window_size) or (shape, window_size, seed) the
        reads are hashed to minimisers instead
        (hibf_gpu_hibf_serve_reads_minimisers) and k is not used.
        With pinned, the bases are placed in pinned host memory first, as a
        server's receive buffer would hold them.  Returns (latency_s,
        ids[nq, R], counts, stats dict)."""
        data, o = reads if isinstance(reads, tuple) else reads_csr(reads)
        o = np.ascontiguousarray(o, dtype=np.uint64)
        if pinned:
            b = torch.empty(max(int(data.size), 1), dtype=torch.uint8, pin_memory=True)
            b.numpy()[: data.size] = data
        else:
            b = np.ascontiguousarray(data if data.size else np.zeros(1, dtype=np.uint8), dtype=np.uint8)
        a = np.ascontiguousarray(arrival_s, dtype=np.float64)
        nq = o.size - 1
        if nq < 0 or a.size != nq or (nq > 0 and int(o[-1]) > data.size):
            raise ValueError("offsets need nq + 1 entries within the bases and arrival_s one per read")
        lat = np.zeros(nq, dtype=np.float64)
        ids = np.zeros((nq, results_per_query), dtype=np.uint64)
        cnt = np.zeros(nq, dtype=np.uint32)
        p = native.ServeParams(window_us, max_batch, max_batch_bases, budget_bytes, streams, results_per_query)
        st = native.ServeStats()
        P = lambda x: ctypes.c_void_p(x.data_ptr()) if isinstance(x, torch.Tensor) \
            else x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        if minimiser is None:
            check(lib().hibf_gpu_hibf_serve_reads(self.h, ctypes.byref(p), P(b), P(o), k, float(fraction), P(a), nq,
                                                  P(lat), P(ids), P(cnt), ctypes.byref(st)))
        else:
            shape, window = minimiser[0], minimiser[1]
            seed = minimiser[2] if len(minimiser) > 2 else MINIMISER_SEED
            check(lib().hibf_gpu_hibf_serve_reads_minimisers(self.h, ctypes.byref(p), P(b), P(o), shape, window, seed,
                                                             float(fraction), P(a), nq, P(lat), P(ids), P(cnt),
                                                             ctypes.byref(st)))
        stats = {k_: getattr(st, k_) for k_, _ in native.ServeStats._fields_}
        return lat, ids, cnt, stats

    def bulk_count_into(self, agent: Agent, values: torch.Tensor, offsets: torch.Tensor, thresholds: torch.Tensor,
                        out: torch.Tensor, counter_bytes: int = 2) -> None:
        """HIBF bulk_count into a preallocated (nq, number_of_user_bins) device
        tensor; thresholds is a uint64 device tensor of one entry per query."""
        nq = offsets.numel() - 1
        _check_buffer(offsets, "offsets", 1)
        _check_buffer(values, "values", 1 if nq else 0)
        if counter_bytes not in _DTYPES:
            raise ValueError("counter_bytes must be 1, 2, 4 or 8")
        _check_buffer(out, "out", nq * self.number_of_user_bins, counter_bytes)
        _check_buffer(thresholds, "thresholds", nq, 8, exact=True)
        check(lib().hibf_gpu_hibf_bulk_count(self.h, agent.h, _ptr(values), _ptr(offsets), nq, _ptr(thresholds),
                                             counter_bytes, _ptr(out), _stream()))

    def bulk_count(self, values, offsets, thresholds, counter_bytes: int = 2, agent: Agent | None = None):
        v = as_device_u64(values)
        o = as_device_u64(offsets)
        nq = o.numel() - 1
        t = as_device_u64(np.broadcast_to(np.asarray(thresholds, dtype=np.uint64), (nq,)))
        agent = agent or self._default_agent()
        out = torch.empty((nq, self.number_of_user_bins), dtype=_DTYPES[counter_bytes], device=o.device)
        check(lib().hibf_gpu_hibf_bulk_count(self.h, agent.h, _ptr(v), _ptr(o), nq, _ptr(t), counter_bytes,
                                             _ptr(out), _stream()))
        return out

    def _default_agent(self) -> Agent:
        if self._agent is None:
            self._agent = Agent()
        return self._agent


_NP_VIEW = {torch.uint8: (torch.uint8, np.uint8), torch.uint16: (torch.int16, np.uint16),
            torch.uint32: (torch.int32, np.uint32), torch.uint64: (torch.int64, np.uint64)}


class Server:
    """The online serving loop (hibf_gpu_server_*): requests are submitted
    while it runs, from any number of threads, and results are collected as
    their batches finish.  The HIBF must outlive the server."""

    def __init__(self, hibf: "DeviceHibf", window_us: float = 200.0, max_batch: int = 65536,
                 max_batch_values: int = 1 << 24, budget_bytes: int = 4 << 30, streams: int = 4,
                 results_per_query: int = 16):
        _dev()
        self._hibf = hibf
        self.results_per_query = int(results_per_query)
        p = native.ServeParams(window_us, max_batch, max_batch_values, budget_bytes, streams, results_per_query)
        h = ctypes.c_void_p()
        check(lib().hibf_gpu_server_create(hibf.h, ctypes.byref(p), ctypes.byref(h)))
        self.h = h

    def submit(self, values, threshold: int) -> int:
        """Queues one request; returns its ticket."""
        v = np.ascontiguousarray(np.asarray(values, dtype=np.uint64))
        ticket = ctypes.c_uint64()
        check(lib().hibf_gpu_server_submit(self.h, v.ctypes.data_as(ctypes.c_void_p), v.size, int(threshold),
                                           ctypes.byref(ticket)))
        return int(ticket.value)

    def poll(self, max_results: int = 4096, wait_s: float = 0.0) -> list:
        """Finished requests as (ticket, ids, count, latency_s); ids holds up to
        results_per_query ids in the reference's depth-first order."""
        R = self.results_per_query
        res = (native.ServerResult * max(max_results, 1))()
        ids = np.zeros((max(max_results, 1), R), dtype=np.uint64)
        n = ctypes.c_uint64()
        check(lib().hibf_gpu_server_poll(self.h, res, ids.ctypes.data_as(ctypes.c_void_p), max_results,
                                         float(wait_s), ctypes.byref(n)))
        return [(int(res[i].ticket), ids[i, :res[i].kept].tolist(), int(res[i].count), float(res[i].latency_s))
                for i in range(n.value)]

    def drain(self, wait_s: float = 60.0) -> None:
        check(lib().hibf_gpu_server_drain(self.h, float(wait_s)))

    def stats(self) -> dict:
        st = native.ServeStats()
        check(lib().hibf_gpu_server_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in native.ServeStats._fields_}

    def close(self) -> None:
        h, self.h = getattr(self, "h", None), None
        if h:
            check(lib().hibf_gpu_server_destroy(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def to_numpy(t: torch.Tensor) -> np.ndarray:
    """Host copy of an unsigned device tensor as the matching numpy dtype."""
    sview, ndt = _NP_VIEW[t.dtype]
    src = t if sview == t.dtype else t.view(sview)
    return src.cpu().numpy().view(ndt)
