"""C-ABI checks that need no GPU: the library loads, exports every function
include/hibf_gpu.h declares, and rejects invalid arguments with the
reference's error categories before touching the device."""
import ctypes

import pytest

from hibf_amd import native


def test_library_exports_every_declared_symbol():
    L = native.lib()
    declared = native.declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(native._SIGNATURES)


@pytest.mark.parametrize("bins,bin_size,h,code", [
    (0, 32, 2, native.ERR_BIN_COUNT),    # "The number of bins must be > 0."
    (64, 32, 0, native.ERR_HASH_COUNT),  # "The number of hash functions must be > 0 and <= 5."
    (64, 32, 6, native.ERR_HASH_COUNT),
    (64, 0, 2, native.ERR_BIN_SIZE),     # "The size of a bin must be > 0."
    (0, 0, 0, native.ERR_BIN_COUNT),     # checked in the reference's order
])
def test_ibf_create_rejects_like_reference(bins, bin_size, h, code):
    L = native.lib()
    out = ctypes.c_void_p()
    assert L.hibf_gpu_ibf_create(bins, bin_size, h, ctypes.byref(out)) == code
    assert out.value is None
    assert L.hibf_gpu_last_error()


def test_null_arguments_are_rejected():
    L = native.lib()
    assert L.hibf_gpu_ibf_create(1, 1, 1, None) == native.ERR_INVALID_ARGUMENT
    assert L.hibf_gpu_ibf_bulk_count(None, None, None, 0, 2, None, None) == native.ERR_INVALID_ARGUMENT
    assert L.hibf_gpu_agent_create(0, None) == native.ERR_INVALID_ARGUMENT
    assert L.hibf_gpu_ibf_destroy(None) == native.OK
    assert L.hibf_gpu_status_string(native.ERR_CAPACITY) == b"output buffer too small"
