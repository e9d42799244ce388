"""ctypes binding of libhibf_gpu.so (the C-ABI declared in include/hibf_gpu.h).

The library is the only compute path: importing works on a machine without a
GPU (for symbol and argument checks), but every query needs the HIP device and
fails loudly if the library or the device is missing.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libhibf_gpu.so")
HEADER = os.path.join(ROOT, "include", "hibf_gpu.h")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_BIN_COUNT = 2
ERR_HASH_COUNT = 3
ERR_BIN_SIZE = 4
ERR_OUT_OF_MEMORY = 5
ERR_DEVICE = 6
ERR_CAPACITY = 7
ERR_UNSUPPORTED = 8
ERR_CONFIG = 9
ERR_BUDGET = 10
ERR_NO_DEVICE = 11
ERR_IO = 12

BIN_MERGED = 0xFFFFFFFFFFFFFFFF
BIN_DELETED = 0xFFFFFFFFFFFFFFFE


class HibfGpuError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"[{code}] {what}")
        self.code = code


class Config(ctypes.Structure):
    """hibf_gpu_config: numeric fields of seqan::hibf::config (config.hpp:111-324)."""

    _fields_ = [
        ("number_of_user_bins", ctypes.c_uint64),
        ("number_of_hash_functions", ctypes.c_uint64),
        ("maximum_fpr", ctypes.c_double),
        ("relaxed_fpr", ctypes.c_double),
        ("threads", ctypes.c_uint64),
        ("sketch_bits", ctypes.c_uint64),
        ("tmax", ctypes.c_uint64),
        ("empty_bin_fraction", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("max_rearrangement_ratio", ctypes.c_double),
        ("disable_estimate_union", ctypes.c_int32),
        ("disable_rearrangement", ctypes.c_int32),
        ("track_occupancy", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServeParams(ctypes.Structure):
    _fields_ = [("window_us", ctypes.c_double), ("max_batch", ctypes.c_uint64), ("max_batch_values", ctypes.c_uint64),
                ("budget_bytes", ctypes.c_uint64), ("streams", ctypes.c_uint32), ("results_per_query", ctypes.c_uint32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("count", ctypes.c_uint32), ("kept", ctypes.c_uint32),
                ("latency_s", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("batches", ctypes.c_uint64), ("lookups", ctypes.c_uint64),
                ("row_bytes", ctypes.c_uint64), ("max_in_flight_bytes", ctypes.c_uint64),
                ("overflowed", ctypes.c_uint64), ("longest_pass_s", ctypes.c_double),
                ("longest_launch_s", ctypes.c_double)]


class Geometry(ctypes.Structure):
    _fields_ = [
        ("bins", ctypes.c_uint64),
        ("technical_bins", ctypes.c_uint64),
        ("bin_size", ctypes.c_uint64),
        ("hash_shift", ctypes.c_uint64),
        ("bin_words", ctypes.c_uint64),
        ("hash_funs", ctypes.c_uint64),
        ("n_words", ctypes.c_uint64),
    ]


_V = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_INT = ctypes.c_int

_SIGNATURES = {
    "hibf_gpu_status_string": (ctypes.c_char_p, [_INT]),
    "hibf_gpu_last_error": (ctypes.c_char_p, []),
    "hibf_gpu_bin_size_in_bits": (_U64, [ctypes.c_double, _U64, _U64]),
    "hibf_gpu_malloc": (_INT, [ctypes.POINTER(_V), _U64]),
    "hibf_gpu_free": (_INT, [_V]),
    "hibf_gpu_memcpy": (_INT, [_V, _V, _U64, _INT]),
    "hibf_gpu_synchronize": (_INT, []),
    "hibf_gpu_tune_set": (_INT, [ctypes.c_char_p, ctypes.c_int64]),
    "hibf_gpu_tune_get": (_INT, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "hibf_gpu_ibf_create": (_INT, [_U64, _U64, _U32, ctypes.POINTER(_V)]),
    "hibf_gpu_ibf_destroy": (_INT, [_V]),
    "hibf_gpu_ibf_get_geometry": (_INT, [_V, ctypes.POINTER(Geometry)]),
    "hibf_gpu_ibf_device_words": (_INT, [_V, ctypes.POINTER(_V)]),
    "hibf_gpu_ibf_upload": (_INT, [_V, _V, _U64]),
    "hibf_gpu_ibf_download": (_INT, [_V, _V, _U64]),
    "hibf_gpu_ibf_emplace": (_INT, [_V, _V, _V, _U64, _U64, _V]),
    "hibf_gpu_ibf_clear": (_INT, [_V, _V, _U64, _V]),
    "hibf_gpu_ibf_emplace_synthetic": (_INT, [_V, _U64, _U64, _U64, _U64, _V]),
    "hibf_gpu_synth_value": (_U64, [_U64, _U64, _U64]),
    "hibf_gpu_ibf_bulk_contains": (_INT, [_V, _V, _U64, _V, _V]),
    "hibf_gpu_ibf_bulk_count": (_INT, [_V, _V, _V, _U64, _U32, _V, _V]),
    "hibf_gpu_agent_create": (_INT, [_U64, ctypes.POINTER(_V)]),
    "hibf_gpu_agent_destroy": (_INT, [_V]),
    "hibf_gpu_agent_stats": (_INT, [_V, ctypes.POINTER(_U64), ctypes.POINTER(_U64), ctypes.POINTER(_U64), _INT]),
    "hibf_gpu_agent_status": (_INT, [_V, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    "hibf_gpu_hibf_membership_prepare": (_INT, [_V, _V, _U64]),
    "hibf_gpu_hibf_membership_launch": (_INT, [_V, _V, _V, _V, _U64, _V, _V, _U64, _V, _V, _V]),
    "hibf_gpu_hibf_membership_launch_noalloc": (_INT, [_V, _V, _V, _U64, _V, _V, _U64, _V, _V, _V]),
    "hibf_gpu_ibf_membership_for": (_INT, [_V, _V, _V, _V, _U64, _V, _V, _U64, _V, ctypes.POINTER(_U64), _V]),
    "hibf_gpu_hibf_create": (_INT, [_U64, ctypes.POINTER(Geometry), _V, _V, _V, _V, _V, _U64, ctypes.POINTER(_V)]),
    "hibf_gpu_hibf_destroy": (_INT, [_V]),
    "hibf_gpu_hibf_emplace": (_INT, [_V, _U64, _V, _V, _U64, _U64, _V]),
    "hibf_gpu_hibf_download_arena": (_INT, [_V, _V, _U64]),
    "hibf_gpu_hibf_arena_words": (_INT, [_V, ctypes.POINTER(_U64)]),
    "hibf_gpu_hibf_device_arena": (_INT, [_V, ctypes.POINTER(_V)]),
    "hibf_gpu_hibf_ibf_view": (_INT, [_V, _U64, ctypes.POINTER(_V)]),
    "hibf_gpu_hibf_build_times": (_INT, [_V, _V]),
    "hibf_gpu_hibf_layout": (_INT, [_V, ctypes.c_char_p, _U64, ctypes.POINTER(_U64)]),
    "hibf_gpu_hibf_layout_times": (_INT, [_V, _V]),
    "hibf_gpu_ibf_occupancy": (_INT, [_V, _V, ctypes.POINTER(_INT)]),
    "hibf_gpu_ibf_try_increase_bin_number_to": (_INT, [_V, _U64, ctypes.POINTER(_INT)]),
    "hibf_gpu_ibf_increase_bin_number_to": (_INT, [_V, _U64]),
    "hibf_gpu_ibf_set_occupancy": (_INT, [_V, _V, _INT]),
    "hibf_gpu_kmer_hash": (_INT, [_V, _V, _U64, _U32, _V, _V, _U64, ctypes.POINTER(_U64), _V]),
    "hibf_gpu_minimiser_hash": (_INT, [_V, _V, _U64, _U64, _U32, _U64, _V, _V, _U64, ctypes.POINTER(_U64), _V]),
    "hibf_gpu_kmer_thresholds": (_INT, [_V, _U64, ctypes.c_double, _V, _V]),
    "hibf_gpu_hll_sketch": (_INT, [_V, _V, _U64, ctypes.c_uint32, _V, _V]),
    "hibf_gpu_hll_sketch_synthetic": (_INT, [_U64, _V, _U64, ctypes.c_uint32, _V, _V]),
    "hibf_gpu_hll_merge": (_INT, [_V, ctypes.c_uint32, _V, _V, _U64, _V, _V]),
    "hibf_gpu_hll_estimate": (_INT, [_V, _U64, ctypes.c_uint32, _V, _V]),
    "hibf_gpu_hll_union_estimates": (_INT, [_V, _U64, ctypes.c_uint32, _V, _U64, _V, _V]),
    "hibf_gpu_ibf_save": (_INT, [_V, ctypes.c_char_p, _INT]),
    "hibf_gpu_ibf_load": (_INT, [ctypes.c_char_p, _U64, ctypes.POINTER(_V), ctypes.POINTER(_U64)]),
    "hibf_gpu_hibf_save": (_INT, [_V, ctypes.c_char_p, _INT]),
    "hibf_gpu_hibf_load": (_INT, [ctypes.c_char_p, _U64, ctypes.POINTER(_V), ctypes.POINTER(_U64)]),
    "hibf_gpu_hibf_membership_for": (_INT, [_V, _V, _V, _V, _U64, _V, _V, _U64, _V, ctypes.POINTER(_U64), _V]),
    "hibf_gpu_hibf_bulk_count": (_INT, [_V, _V, _V, _V, _U64, _V, _U32, _V, _V]),
    "hibf_gpu_hibf_serve": (_INT, [_V, ctypes.POINTER(ServeParams), _V, _V, _V, _V, _U64, _V, _V, _V,
                                   ctypes.POINTER(ServeStats)]),
    "hibf_gpu_hibf_serve_reads": (_INT, [_V, ctypes.POINTER(ServeParams), _V, _V, ctypes.c_uint32, ctypes.c_double,
                                         _V, _U64, _V, _V, _V, ctypes.POINTER(ServeStats)]),
    "hibf_gpu_hibf_serve_reads_minimisers": (_INT, [_V, ctypes.POINTER(ServeParams), _V, _V, _U64, _U32, _U64,
                                                    ctypes.c_double, _V, _U64, _V, _V, _V,
                                                    ctypes.POINTER(ServeStats)]),
    "hibf_gpu_server_create": (_INT, [_V, ctypes.POINTER(ServeParams), ctypes.POINTER(_V)]),
    "hibf_gpu_server_submit": (_INT, [_V, _V, _U64, ctypes.c_uint16, ctypes.POINTER(_U64)]),
    "hibf_gpu_server_poll": (_INT, [_V, _V, _V, _U64, ctypes.c_double, ctypes.POINTER(_U64)]),
    "hibf_gpu_server_drain": (_INT, [_V, ctypes.c_double]),
    "hibf_gpu_server_stats": (_INT, [_V, ctypes.POINTER(ServeStats)]),
    "hibf_gpu_server_destroy": (_INT, [_V]),
    "hibf_gpu_hibf_info": (_INT, [_V] + [ctypes.POINTER(_U64)] * 5),
    "hibf_gpu_hibf_export_tables": (_INT, [_V, ctypes.POINTER(Geometry), _V, _V, _V, _V]),
    "hibf_gpu_config_default": (_INT, [ctypes.POINTER(Config)]),
    "hibf_gpu_config_validate": (_INT, [ctypes.POINTER(Config)]),
    "hibf_gpu_layout_compute": (_INT, [ctypes.POINTER(Config), _V, ctypes.c_char_p, _U64, ctypes.POINTER(_U64)]),
    "hibf_gpu_layout_compute_sketched": (_INT, [ctypes.POINTER(Config), _V, _V, ctypes.c_char_p, _U64,
                                                ctypes.POINTER(_U64), _V]),
    "hibf_gpu_layout_compute_host_sketches": (_INT, [ctypes.POINTER(Config), _V, _V, _V, _U64, ctypes.c_char_p, _U64,
                                                     ctypes.POINTER(_U64), _V]),
    "hibf_gpu_hibf_build": (_INT, [ctypes.POINTER(Config), _V, _V, ctypes.c_char_p, ctypes.POINTER(_V)]),
    "hibf_gpu_hibf_build_synthetic": (_INT, [ctypes.POINTER(Config), _V, _U64, ctypes.c_char_p, ctypes.POINTER(_V)]),
    "hibf_gpu_hibf_plan_stats": (_INT, [ctypes.POINTER(Config), _V, ctypes.c_char_p] + [ctypes.POINTER(_U64)] * 4),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libhibf_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hibf_amd.build`")
        # torch's wheel carries its own HIP runtime (libamdhip64.so, soname
        # libamdhip64.so.7).  Loading torch first makes this library's
        # libamdhip64.so.7 dependency resolve to that same runtime; loaded the
        # other way round, the process would hold two HIP/HSA runtimes and the
        # second to initialise finds no device.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_functions() -> list[str]:
    """Names of the functions declared in include/hibf_gpu.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|uint64_t|const char \*)\s+(hibf_gpu_\w+)\s*\(", text, re.M)))


def tune_set(key: str, value: int) -> None:
    check(lib().hibf_gpu_tune_set(key.encode(), int(value)))


def tune_get(key: str) -> int:
    v = ctypes.c_int64()
    check(lib().hibf_gpu_tune_get(key.encode(), ctypes.byref(v)))
    return int(v.value)


def make_config(**kw) -> Config:
    """A hibf_gpu_config with the reference defaults, overridden by keywords."""
    c = Config()
    check(lib().hibf_gpu_config_default(ctypes.byref(c)))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(f"unknown config field {k}")
        setattr(c, k, int(v) if isinstance(v, bool) else v)
    return c


def validate_config(c: Config) -> Config:
    check(lib().hibf_gpu_config_validate(ctypes.byref(c)))
    return c


def compute_layout(c: Config, kmer_counts) -> str:
    """Layout text (reference format) for the given user-bin cardinalities."""
    import numpy as np
    k = np.ascontiguousarray(np.asarray(kmer_counts, dtype=np.uint64))
    need = ctypes.c_uint64(0)
    rc = lib().hibf_gpu_layout_compute(ctypes.byref(c), k.ctypes.data_as(ctypes.c_void_p), None, 0,
                                       ctypes.byref(need))
    if rc != ERR_CAPACITY:
        check(rc)
    buf = ctypes.create_string_buffer(need.value)
    check(lib().hibf_gpu_layout_compute(ctypes.byref(c), k.ctypes.data_as(ctypes.c_void_p), buf, need.value,
                                        ctypes.byref(need)))
    return buf.value.decode()


def compute_layout_sketched(c: Config, kmer_counts, d_registers: int, times: list | None = None) -> str:
    """Layout text with union estimates and rearrangement from the user bins'
    sketches (device address of number_of_user_bins x 2^sketch_bits registers).
    `times`, if given, receives [union estimation s, rearrangement s]."""
    import numpy as np
    k = np.ascontiguousarray(np.asarray(kmer_counts, dtype=np.uint64))
    kp = k.ctypes.data_as(ctypes.c_void_p)
    t = (ctypes.c_double * 2)()
    need = ctypes.c_uint64(0)
    buf = ctypes.create_string_buffer(16 << 20)  # the text of 32,768 user bins is about 1.5 MB
    rc = lib().hibf_gpu_layout_compute_sketched(ctypes.byref(c), kp, d_registers, buf, len(buf), ctypes.byref(need),
                                                ctypes.cast(t, ctypes.c_void_p))
    if rc == ERR_CAPACITY:  # a second pass with room for the text
        buf = ctypes.create_string_buffer(need.value)
        rc = lib().hibf_gpu_layout_compute_sketched(ctypes.byref(c), kp, d_registers, buf, need.value,
                                                    ctypes.byref(need), ctypes.cast(t, ctypes.c_void_p))
    check(rc)
    if times is not None:
        times[:] = [t[0], t[1]]
    return buf.value.decode()


def plan_stats(c: Config, sizes, layout_text: str | None = None) -> dict:
    """Host-only dry run of a build: IBFs, arena words, segments, insertions."""
    import numpy as np
    sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.uint64))
    out = [ctypes.c_uint64() for _ in range(4)]
    check(lib().hibf_gpu_hibf_plan_stats(ctypes.byref(c), sz.ctypes.data_as(ctypes.c_void_p),
                                         None if layout_text is None else layout_text.encode(),
                                         *[ctypes.byref(o) for o in out]))
    return dict(zip(("n_ibfs", "n_words", "n_segments", "n_emplaces"), (o.value for o in out)))


def check(rc: int) -> None:
    if rc != OK:
        L = lib()
        raise HibfGpuError(rc, f"{L.hibf_gpu_status_string(rc).decode()}: {L.hibf_gpu_last_error().decode()}")
