"""Archive layout and archive error handling without a GPU: the byte offsets
of the cereal binary layout, and malformed files rejected with
HIBF_GPU_ERR_IO before any device memory is touched."""
import struct

import numpy as np
import pytest

import cereal_layout as CL

HIBF_GPU_ERR_IO = 12


def test_ibf_layout_offsets():
    """Field positions of interleaved_bloom_filter::serialize in a binary archive."""
    words = np.arange(5, dtype=np.uint64) * np.uint64(0x0101010101010101)
    occ = np.arange(64, dtype=np.uint64)
    b = CL.ibf_bytes(CL.ibf_record(3, 5, 2, words, occ, True))
    assert len(b) == 4 + 7 * 8 + 8 * 5 + 8 + 8 * 64 + 1
    assert b[:4] == b"\x01\x00\x00\x00"                                     # version 1
    assert struct.unpack_from("<7Q", b, 4) == (3, 64, 5, 61, 1, 2, 320)    # bins .. hash_funs, bit count
    assert np.frombuffer(b, "<u8", 5, 60).tolist() == words[:5].tolist()
    assert struct.unpack_from("<Q", b, 100)[0] == 64                        # occupancy length
    assert b[-1:] == b"\x01"                                               # track_occupancy
    rec, end = CL.parse_ibf(b)
    assert end == len(b) and rec.hash_shift == 61 and rec.occupancy.tolist() == occ.tolist()


def test_hibf_layout_offsets():
    ibf = CL.ibf_record(1, 1, 1, np.zeros(1, dtype=np.uint64))
    r = CL.HibfRecord(7, [ibf], [np.array([0], dtype=np.uint64)], [(0, 0)], [np.array([6], dtype=np.uint64)])
    b = CL.hibf_bytes(r)
    assert struct.unpack_from("<IQQ", b, 0) == (1, 7, 1)
    n_ibf = len(CL.ibf_bytes(ibf))
    tail = b[20 + n_ibf:]
    assert struct.unpack("<QQQ", tail[:24]) == (1, 1, 0)        # next_ibf_id: 1 IBF, 1 entry, id 0
    assert struct.unpack("<QQQ", tail[24:48]) == (1, 0, 0)      # prev_ibf_id: 1 pair (0, 0)
    assert struct.unpack("<QQQ", tail[48:72]) == (1, 1, 6)      # ibf_bin_to_user_bin_id
    back, end = CL.parse_hibf(b)
    assert end == len(b) and back.number_of_user_bins == 7


def _load_status(fn, path, offset=0):
    import ctypes
    from hibf_amd import native
    h, end = ctypes.c_void_p(), ctypes.c_uint64()
    rc = getattr(native.lib(), fn)(str(path).encode(), offset, ctypes.byref(h), ctypes.byref(end))
    return rc, native.lib().hibf_gpu_last_error().decode()


def _ibf(bins=70, bin_size=9):
    rec = CL.ibf_record(bins, bin_size, 2, np.arange(((bins + 63) // 64) * bin_size, dtype=np.uint64))
    return rec


@pytest.mark.parametrize("cut", [3, 30, 70, -9, -1])
def test_truncated_ibf_archive(tmp_path, cut):
    b = CL.ibf_bytes(_ibf())
    p = tmp_path / "ibf.bin"
    p.write_bytes(b[:cut])
    rc, msg = _load_status("hibf_gpu_ibf_load", p)
    assert rc == HIBF_GPU_ERR_IO and "ends early" in msg


@pytest.mark.parametrize("field,value", [("technical_bins", 65), ("bin_words", 3), ("hash_shift", 0),
                                         ("hash_funs", 6), ("bits", 64), ("bins", 0), ("version", 2)])
def test_inconsistent_ibf_archive(tmp_path, field, value):
    rec = _ibf()
    setattr(rec, field, value)
    p = tmp_path / "ibf.bin"
    p.write_bytes(CL.ibf_bytes(rec))
    rc, msg = _load_status("hibf_gpu_ibf_load", p)
    assert rc == HIBF_GPU_ERR_IO, msg


def test_archive_offset_and_missing_file(tmp_path):
    p = tmp_path / "ibf.bin"
    p.write_bytes(CL.ibf_bytes(_ibf()))
    assert _load_status("hibf_gpu_ibf_load", p, p.stat().st_size + 1)[0] == HIBF_GPU_ERR_IO
    assert _load_status("hibf_gpu_ibf_load", tmp_path / "absent.bin")[0] == HIBF_GPU_ERR_IO
    assert _load_status("hibf_gpu_hibf_load", tmp_path / "absent.bin")[0] == HIBF_GPU_ERR_IO


def test_malformed_hibf_archives(tmp_path):
    ibf = _ibf(3, 4)
    good = CL.HibfRecord(3, [ibf], [np.zeros(3, dtype=np.uint64)], [(0, 0)], [np.arange(3, dtype=np.uint64)])
    cases = {
        "version": CL.HibfRecord(3, good.ibfs, good.next_ibf_id, good.prev_ibf_id, good.ibf_bin_to_user_bin_id,
                                 version=0),
        "next count": CL.HibfRecord(3, good.ibfs, good.next_ibf_id * 2, good.prev_ibf_id,
                                    good.ibf_bin_to_user_bin_id),
        "prev count": CL.HibfRecord(3, good.ibfs, good.next_ibf_id, [], good.ibf_bin_to_user_bin_id),
        "table length": CL.HibfRecord(3, good.ibfs, [np.zeros(2, dtype=np.uint64)], good.prev_ibf_id,
                                      good.ibf_bin_to_user_bin_id),
        "no ibf": CL.HibfRecord(3, [], [], [], []),
    }
    for name, rec in cases.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(CL.hibf_bytes(rec))
        rc, msg = _load_status("hibf_gpu_hibf_load", p)
        assert rc == HIBF_GPU_ERR_IO, (name, msg)
    b = CL.hibf_bytes(good)
    for cut in (10, 25, len(b) - 5):
        p = tmp_path / "cut.bin"
        p.write_bytes(b[:cut])
        assert _load_status("hibf_gpu_hibf_load", p)[0] == HIBF_GPU_ERR_IO, cut
