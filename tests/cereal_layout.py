"""Byte layout of a cereal::BinaryOutputArchive of the reference's
interleaved_bloom_filter and hierarchical_interleaved_bloom_filter (test
infrastructure).

Written from cereal's binary-archive rules: an arithmetic value is its
little-endian bytes; a std::vector of an arithmetic type is a uint64 count and
then the elements; a std::vector of a class is the uint64 count and then each
element's fields; binary_data is raw bytes; a base class's fields come in
place; make_nvp stores the value alone; there are no version records because
no serialize function on this path takes a version argument.  Fields and their
order follow interleaved_bloom_filter.hpp:471-485, bit_vector.hpp:885-924 and
hierarchical_interleaved_bloom_filter.hpp:267-277, 300-309.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

IBF_HEADER_BYTES = 4 + 7 * 8  # version, six geometry fields, bit count


@dataclass
class IbfRecord:
    bins: int
    technical_bins: int
    bin_size: int
    hash_shift: int
    bin_words: int
    hash_funs: int
    bits: int
    words: np.ndarray
    occupancy: np.ndarray
    track_occupancy: bool
    version: int = 1


def ibf_record(bins: int, bin_size: int, hash_funs: int, words: np.ndarray, occupancy=None,
               track_occupancy: bool = False) -> IbfRecord:
    """The record of an IBF with geometry as src/interleaved_bloom_filter.cpp:48-52 derives it."""
    bin_words = (bins + 63) // 64
    tb = bin_words * 64
    occ = np.zeros(tb, dtype=np.uint64) if occupancy is None else np.asarray(occupancy, dtype=np.uint64)
    return IbfRecord(bins, tb, bin_size, 64 - bin_size.bit_length(), bin_words, hash_funs, tb * bin_size,
                     np.asarray(words, dtype=np.uint64), occ, track_occupancy)


def ibf_bytes(r: IbfRecord) -> bytes:
    out = [struct.pack("<I7Q", r.version, r.bins, r.technical_bins, r.bin_size, r.hash_shift, r.bin_words,
                       r.hash_funs, r.bits)]
    out.append(np.ascontiguousarray(r.words, dtype="<u8").tobytes())
    out.append(struct.pack("<Q", r.occupancy.size))
    out.append(np.ascontiguousarray(r.occupancy, dtype="<u8").tobytes())
    out.append(struct.pack("<?", r.track_occupancy))
    return b"".join(out)


def parse_ibf(buf: bytes, pos: int = 0) -> tuple[IbfRecord, int]:
    version, bins, tb, bin_size, shift, bw, h, bits = struct.unpack_from("<I7Q", buf, pos)
    pos += IBF_HEADER_BYTES
    n_words = (bits + 63) // 64
    words = np.frombuffer(buf, dtype="<u8", count=n_words, offset=pos).astype(np.uint64)
    pos += 8 * n_words
    (n_occ,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    occ = np.frombuffer(buf, dtype="<u8", count=n_occ, offset=pos).astype(np.uint64)
    pos += 8 * n_occ
    (track,) = struct.unpack_from("<?", buf, pos)
    pos += 1
    return IbfRecord(bins, tb, bin_size, shift, bw, h, bits, words, occ, bool(track), version), pos


@dataclass
class HibfRecord:
    number_of_user_bins: int
    ibfs: list = field(default_factory=list)
    next_ibf_id: list = field(default_factory=list)
    prev_ibf_id: list = field(default_factory=list)   # (ibf_idx, bin_idx) per IBF
    ibf_bin_to_user_bin_id: list = field(default_factory=list)
    version: int = 1


def _nested_bytes(rows) -> bytes:
    out = [struct.pack("<Q", len(rows))]
    for row in rows:
        a = np.ascontiguousarray(row, dtype="<u8")
        out.append(struct.pack("<Q", a.size))
        out.append(a.tobytes())
    return b"".join(out)


def _parse_nested(buf: bytes, pos: int):
    (n,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    rows = []
    for _ in range(n):
        (m,) = struct.unpack_from("<Q", buf, pos)
        pos += 8
        rows.append(np.frombuffer(buf, dtype="<u8", count=m, offset=pos).astype(np.uint64))
        pos += 8 * m
    return rows, pos


def hibf_bytes(r: HibfRecord) -> bytes:
    out = [struct.pack("<IQQ", r.version, r.number_of_user_bins, len(r.ibfs))]
    out += [ibf_bytes(i) for i in r.ibfs]
    out.append(_nested_bytes(r.next_ibf_id))
    out.append(struct.pack("<Q", len(r.prev_ibf_id)))
    out += [struct.pack("<QQ", a, b) for a, b in r.prev_ibf_id]
    out.append(_nested_bytes(r.ibf_bin_to_user_bin_id))
    return b"".join(out)


def parse_hibf(buf: bytes, pos: int = 0) -> tuple[HibfRecord, int]:
    version, n_ub, n = struct.unpack_from("<IQQ", buf, pos)
    pos += 20
    r = HibfRecord(n_ub, version=version)
    for _ in range(n):
        rec, pos = parse_ibf(buf, pos)
        r.ibfs.append(rec)
    r.next_ibf_id, pos = _parse_nested(buf, pos)
    (n_prev,) = struct.unpack_from("<Q", buf, pos)
    pos += 8
    for _ in range(n_prev):
        r.prev_ibf_id.append(struct.unpack_from("<QQ", buf, pos))
        pos += 16
    r.ibf_bin_to_user_bin_id, pos = _parse_nested(buf, pos)
    return r, pos
