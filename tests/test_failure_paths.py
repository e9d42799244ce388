"""Failure paths that need no GPU: malformed layouts are rejected by the host
build plan before any table is sized from them (hibf_gpu_hibf_plan_stats runs
the same plan as the device builds, without device memory), and the C++
batched agents reject a threshold count that differs from the query count
before they touch the device."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from hibf_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEAD = "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS\n"
TOP0 = "#TOP_LEVEL_IBF fullest_technical_bin_idx:0\n"
REST = "1\t1\t1\n2\t2\t1\n3\t3\t1\n"


def _plan(text, n=4, tmax=64):
    c = native.make_config(number_of_user_bins=n, tmax=tmax)
    sizes = np.full(n, 100, dtype=np.uint64)
    out = [ctypes.c_uint64() for _ in range(4)]
    rc = native.lib().hibf_gpu_hibf_plan_stats(ctypes.byref(c), sizes.ctypes.data_as(ctypes.c_void_p),
                                               text.encode(), *[ctypes.byref(o) for o in out])
    return rc, native.lib().hibf_gpu_last_error().decode(), [o.value for o in out]


def test_well_formed_layout_plans():
    rc, _, (n_ibfs, _, n_segments, n_emplaces) = _plan(TOP0 + HEAD + "0\t0\t1\n" + REST)
    assert rc == native.OK and n_ibfs == 1 and n_segments == 4 and n_emplaces == 400


@pytest.mark.parametrize("text, message", [
    # a merged bin whose lower-level IBF holds no user bin (its index is past the parent's bins)
    (TOP0 + "#LOWER_LEVEL_IBF_5 fullest_technical_bin_idx:0\n" + HEAD + "0\t0\t1\n" + REST, "merged bin 5"),
    # the fullest technical bin past the IBF's bins
    ("#TOP_LEVEL_IBF fullest_technical_bin_idx:9\n" + HEAD + "0\t0\t1\n" + REST, "fullest technical bin 9"),
    # a lower-level user bin split over more technical bins than the top level's correction table covers
    (TOP0 + "#LOWER_LEVEL_IBF_0 fullest_technical_bin_idx:0\n" + HEAD + "0\t0;0\t1;100\n" + REST, "split over 100"),
    # technical-bin indices that would size tables of billions of entries
    (TOP0 + HEAD + "0\t4294967296\t1\n" + REST, "technical bin index beyond"),
    (TOP0 + "#LOWER_LEVEL_IBF_0 fullest_technical_bin_idx:0\n" + HEAD + "0\t0;99999999999\t1;1\n" + REST,
     "technical bin index beyond"),
    # a user bin that does not exist, one that occupies no technical bin
    (TOP0 + HEAD + "9\t0\t1\n" + REST, "user bin index out of range"),
    (TOP0 + HEAD + "0\t0\t0\n" + REST, "occupies no technical bin"),
    # a user bin below a merged bin that has no lower-level IBF
    (TOP0 + HEAD + "0\t0;1\t1;1\n" + REST, "without lower-level IBF"),
    # a user bin placed twice
    (TOP0 + HEAD + "0\t0\t1\n0\t1\t1\n2\t2\t1\n3\t3\t1\n", "exactly once"),
    ("garbage", "could not be parsed"),
    ("", "could not be parsed"),
])
def test_malformed_layouts_are_rejected(text, message):
    rc, err, _ = _plan(text)
    assert rc == native.ERR_INVALID_ARGUMENT
    assert message in err, err


def test_cpp_batched_agents_reject_threshold_count_mismatch(tmp_path):
    src = tmp_path / "mismatch.cpp"
    src.write_text(r'''
#include <hibf/hierarchical_interleaved_bloom_filter.hpp>
#include <cstdio>
#include <stdexcept>
#include <vector>
int main()
{
    std::vector<std::vector<uint64_t>> queries{{1, 2, 3}, {4, 5}};
    int caught = 0;
    seqan::hibf::hierarchical_interleaved_bloom_filter::membership_agent_type m{};
    try { (void)m.membership_for_batch(queries, std::vector<uint16_t>{1}); }
    catch (std::invalid_argument const & e) { ++caught; std::puts(e.what()); }
    seqan::hibf::hierarchical_interleaved_bloom_filter::counting_agent_type<uint16_t> c{};
    try { (void)c.bulk_count_batch(queries, std::vector<uint64_t>{1, 2, 3}); }
    catch (std::invalid_argument const & e) { ++caught; std::puts(e.what()); }
    std::printf("caught %d\n", caught);
    return caught == 2 ? 0 : 1;
}
''')
    exe = tmp_path / "mismatch"
    lib_dir = os.path.join(ROOT, "hibf_amd")
    r = subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT}/include", str(src), f"-L{lib_dir}", "-lhibf_gpu",
                        f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "1 thresholds for 2 queries" in r.stdout and "3 thresholds for 2 queries" in r.stdout
