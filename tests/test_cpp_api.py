"""The seqan::hibf C++ headers (include/hibf/) as a drop-in: the reference's
snippet programs, written against them, print the reference's committed
.out files."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "snippets")
EXPECTED = os.path.join(ROOT, "tests", "golden", "cpp_snippets.out")


def test_cpp_headers_compile(tmp_path):
    """Headers compile standalone with the host compiler (no HIP headers needed)."""
    for hdr in ("hibf/interleaved_bloom_filter.hpp", "hibf/hierarchical_interleaved_bloom_filter.hpp",
                "hibf/config.hpp", "hibf/misc/counting_vector.hpp", "hibf/misc/bit_vector.hpp",
                "hibf/misc/print.hpp", "hibf/misc/timer.hpp", "hibf/layout/layout.hpp", "hibf/sketch/hyperloglog.hpp",
                "hibf/sketch/compute_sketches.hpp", "hibf/sketch/toolbox.hpp", "hibf/layout/compute_layout.hpp",
                "hibf_gpu.h"):
        src = tmp_path / "t.cpp"
        src.write_text(f"#include <{hdr}>\nint main() {{}}\n")
        r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", f"-I{ROOT}/include", str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_snippets_print_reference_outputs():
    assert os.path.exists(EXE), "build tests/cpp/snippets first (python -m hibf_amd.build)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = [line.rstrip() for line in r.stdout.splitlines()]
    want = [line.rstrip() for line in open(EXPECTED).read().splitlines()]
    assert got == want


@pytest.mark.gpu
def test_cpp_store_and_load(tmp_path):
    """store()/load() of the C++ classes: loaded objects answer as the originals."""
    exe = os.path.join(ROOT, "tests", "cpp", "archive_roundtrip")
    assert os.path.exists(exe), "build tests/cpp/archive_roundtrip first (python -m hibf_amd.build)"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == 6 and all(line.endswith(" same") for line in lines), r.stdout


@pytest.mark.gpu
def test_cpp_hibf_public_members():
    """ibf_vector (views of the arena), prev_ibf_id, version, the build
    timers and copy semantics of the C++ HIBF class, the reference's
    build_from_layout test (config and layout read from one stream) and its
    emplace_with_occupancy test on the C++ IBF."""
    exe = os.path.join(ROOT, "tests", "cpp", "hibf_members")
    assert os.path.exists(exe), "build tests/cpp/hibf_members first (python -m hibf_amd.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == 9 and all(line.endswith(" ok") for line in lines), r.stdout


def _config_io(tmp_path):
    exe = tmp_path / "config_io"
    r = subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT}/include", os.path.join(ROOT, "tests", "cpp", "config_io.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def _read(exe, text):
    r = subprocess.run([exe, "read"], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    fields, rest = {}, []
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        if k == "rest":
            rest.append(v)
        else:
            fields[k] = v
    return fields, rest


def _want(d):
    return {k: float(v) for k, v in d.items()}


def _num(fields):
    return {k: float(v) for k, v in fields.items()}


def test_config_write_to_and_read_from(tmp_path):
    """config::write_to prints the reference's expected block byte for byte
    and read_from parses versions 1-3 as the reference's config_test.cpp
    expects (version 1: no empty_bin_fraction; version 2: track_occupancy
    follows empty_bin_fraction != 0), leaving the stream after the block."""
    import json
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "config_header.json")))
    exe = _config_io(tmp_path)
    args = [f"{k}={str(v).lower() if isinstance(v, bool) else v}" for k, v in g["write_to_config"].items()]
    r = subprocess.run([exe, "write", *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout == g["write_to_expected"]

    want = _want(g["read_from_expected"])
    got, rest = _read(exe, "@blah some chopper stuff\n" + g["read_from_input"] + "#TOP_LEVEL_IBF fullest_technical_bin_idx:0\n")
    assert _num(got) == want and rest == ["#TOP_LEVEL_IBF fullest_technical_bin_idx:0"]

    v2 = g["read_from_input"].replace('"version": 3', '"version": 2').replace('@        "track_occupancy": true,\n', "")
    got, _ = _read(exe, v2)
    assert _num(got) == want   # track_occupancy from empty_bin_fraction 0.5

    v1 = v2.replace('"version": 2', '"version": 1').replace('@        "empty_bin_fraction": 0.5,\n', "")
    got, _ = _read(exe, v1)
    assert _num(got) == dict(want, empty_bin_fraction=0.0, track_occupancy=0.0)

    # write_to -> read_from round trip keeps every field
    r = subprocess.run([exe, "write", *args, "empty_bin_fraction=0.25", "track_occupancy=true"], capture_output=True,
                       text=True, timeout=60)
    got, _ = _read(exe, r.stdout)
    assert _num(got) == dict(_want(g["write_to_config"]), relaxed_fpr=0.3, empty_bin_fraction=0.25,
                             track_occupancy=1.0)


LAYOUT_FILE = ("#TOP_LEVEL_IBF fullest_technical_bin_idx:111\n"
               "#LOWER_LEVEL_IBF_0 fullest_technical_bin_idx:0\n"
               "#LOWER_LEVEL_IBF_2 fullest_technical_bin_idx:2\n"
               "#LOWER_LEVEL_IBF_1;2;3;4 fullest_technical_bin_idx:22\n"
               "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS\n"
               "7\t0\t1\n"
               "4\t1;0\t1;22\n"
               "5\t1;2;3;4;22\t1;1;1;1;21\n")   # test/unit/hibf/layout/layout_test.cpp:59-68


def test_layout_records_read_and_write(tmp_path):
    """layout::layout read_from/write_to/clear against the reference's
    layout_test.cpp expectations: the records of its layout file, the same text
    written back, and a partitioned file (three layouts in a row, the second
    and third with another fullest bin of LOWER_LEVEL_IBF_0) read one by one."""
    exe = tmp_path / "layout_io"
    r = subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT}/include", os.path.join(ROOT, "tests", "cpp", "layout_io.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    parts = [LAYOUT_FILE.replace("IBF_0 fullest_technical_bin_idx:0", f"IBF_0 fullest_technical_bin_idx:{i}")
             for i in range(3)]
    r = subprocess.run([str(exe)], input="".join(parts), capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    for i, part in enumerate(parts):
        records = (f"top 111\nmax 0 {i}\nmax 2 2\nmax 1;2;3;4 22\n"
                   "ub 7 - 0 1\nub 4 1 0 22\nub 5 1;2;3;4 22 21\n")
        block = records + part + "cleared 1 0\n"
        assert out.startswith(block), (i, out[:400])
        out = out[len(block):]
    assert out == ""


def test_cpp_hyperloglog_matches_oracle(tmp_path):
    """The host seqan::hibf::sketch::hyperloglog class gives the oracle's
    registers and, bit for bit, its estimates (which the reference's HLL test
    expectations pin), including the 7.899522493 known answer."""
    import numpy as np
    import oracle as O
    exe = tmp_path / "hll_host"
    r = subprocess.run(["g++", "-std=c++20", "-O2", f"-I{ROOT}/include", os.path.join(ROOT, "tests", "cpp", "hll_host.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(2)
    small = [255881241332063154, 13132817195163223578, 5120631300412165844, 16862690419523416066,
             148518882728022940, 15892358469365346306, 10885195586503739779, 9563173945158404745]
    cases = [(5, small), (5, []), (6, O.mt19937_64(0, 1500).tolist())]
    for bits in (7, 12, 14):
        for n in (1, 100, 5000, 60000):
            cases.append((bits, rng.integers(0, 2**63, size=n, dtype=np.int64).astype(np.uint64).tolist()))
    text = "".join(f"{b} {len(v)} {' '.join(str(int(x)) for x in v)}\n" for b, v in cases)
    r = subprocess.run([str(exe)], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == len(cases)
    for (bits, v), line in zip(cases, lines):
        regs_hex, est = line.split()
        want = O.hll_sketch(np.asarray(v, dtype=np.uint64), bits)
        assert regs_hex == want.tobytes().hex()
        assert float(est) == O.hll_estimate(want)
    assert abs(float(lines[0].split()[1]) - 7.899522493) < 1e-7


@pytest.mark.gpu
def test_cpp_compute_sketches_on_device():
    """seqan::hibf::sketch::compute_sketches (device) equals the host class's add()."""
    import numpy as np
    exe = os.path.join(ROOT, "tests", "cpp", "hll_device")
    assert os.path.exists(exe), "build tests/cpp/hll_device first (python -m hibf_amd.build)"
    rng = np.random.default_rng(4)
    sizes = [0, 1, 50, 70000, 1000, 3]
    text = "".join(f"12 {n} {' '.join(str(int(x)) for x in rng.integers(0, 2**63, size=n, dtype=np.int64))}\n"
                   for n in sizes)
    r = subprocess.run([exe, "device"], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines()[-1] == f"device same {len(sizes)}"


def test_cpp_sketch_toolbox_known_answers(tmp_path):
    """seqan::hibf::sketch::toolbox (include/hibf/sketch/toolbox.hpp) prints
    every expectation of the reference's toolbox_test.cpp: union estimates of
    the four-sketch fixture (1635 = estimate of 1,500 mt19937_64(0) values),
    the seeded shuffle, prune, rotate, trace, cluster_bins and the two
    rearrange_bins cases (libstdc++ results)."""
    exe = tmp_path / "toolbox"
    r = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", f"-I{ROOT}/include",
                        os.path.join(ROOT, "tests", "cpp", "toolbox.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert lines[-1] == "ok" and len(lines) == 19, r.stdout


@pytest.mark.parametrize("n,tmax,seed", [(500, 64, 1), (3000, 64, 2), (5000, 128, 3), (70, 64, 4)])
def test_layout_union_mode_reduces_to_sum_mode(tmp_path, n, tmax, seed):
    """With union estimates equal to the exact sums of the counts, the DP's
    union-estimate mode places every user bin as the disable_estimate_union
    mode (same user-bin records and merged-bin paths), over several levels."""
    exe = tmp_path / "layout_union"
    r = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", f"-I{ROOT}/include", f"-I{ROOT}/hibf_amd/csrc",
                        os.path.join(ROOT, "tests", "cpp", "layout_union.cpp"),
                        os.path.join(ROOT, "hibf_amd", "csrc", "layout.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(n), str(tmax), str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok "), r.stdout


@pytest.mark.gpu
def test_cpp_compute_layout_on_device():
    """seqan::hibf::layout::compute_layout (union estimates and rearrangement on
    the device): the reference's dispatch case, a subset of positions, the
    timers, and the 300-user-bin layout equal to the library's own sketched
    layout of the same user bins."""
    import numpy as np
    import oracle as O
    from hibf_amd import native
    from hibf_amd.device import hll_estimates, hll_sketches
    exe = os.path.join(ROOT, "tests", "cpp", "compute_layout")
    assert os.path.exists(exe), "build tests/cpp/compute_layout first (python -m hibf_amd.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[:3] == ["dispatch ok", "timers ok", "positions_subset ok"], lines[:3]
    stream = O.mt19937_64(5, 300 * 3100).tolist()
    sets, at = [], 0
    for _ in range(300):
        n = 50 + stream[at] % 3000
        sets.append(np.asarray(stream[at + 1:at + 1 + n], dtype=np.uint64))
        at += 1 + n
    regs = hll_sketches(sets)
    counts = hll_estimates(regs).astype(np.uint64)
    c = native.validate_config(native.make_config(number_of_user_bins=300, tmax=64))
    want = native.compute_layout_sketched(c, counts, regs.data_ptr())
    assert "\n".join(lines[3:]) + "\n" == want


def test_layout_dp_reference_expectations(tmp_path):
    """The layout DP reproduces every expectation of the reference's
    hierarchical_binning_test.cpp (eight trees: split and merged bins, empty-bin
    fraction, alpha 1, two to four levels, merged technical bin 0) and
    simple_binning_test.cpp (two layouts and the too-few-bins error), and the
    FPR corrections of fpr_correction_test.cpp."""
    exe = tmp_path / "binning_goldens"
    r = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", f"-I{ROOT}/include", f"-I{ROOT}/hibf_amd/csrc",
                        os.path.join(ROOT, "tests", "cpp", "binning_goldens.cpp"),
                        os.path.join(ROOT, "hibf_amd", "csrc", "layout.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    lines = r.stdout.splitlines()
    assert r.returncode == 0 and len(lines) == 12 and all(line.endswith(" ok") for line in lines), r.stdout
