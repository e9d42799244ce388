"""The host code that parses untrusted input, under AddressSanitizer and
UndefinedBehaviorSanitizer, on the CPU: layout text and the build plan
(hibf_amd/csrc/layout.cpp, build.cpp, with the replay of the reference's
k-mer sets), config and layout files (include/hibf/config.hpp,
include/hibf/layout/layout.hpp) and archives (hibf_amd/csrc/archive.cpp).
The device side of the C-ABI is replaced by tests/cpp/no_device.cpp, which
reports that there is no GPU, so archive loading stops after parsing.  Inputs:
well-formed files, the malformed layouts of tests/test_failure_paths.py and
seeded mutations (truncations, byte flips, huge numbers, dropped and repeated
lines) of each."""
import json
import os
import random
import subprocess

import numpy as np
import pytest

import cereal_layout as CL
from test_failure_paths import HEAD, REST, TOP0

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hibf_amd", "csrc")
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = tmp_path_factory.mktemp("san") / "host_sanitize"
    cmd = ["g++", "-std=c++20", "-O1", "-g", *SANITIZE, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{ROOT}/include", f"-I{CSRC}",
           os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"), os.path.join(ROOT, "tests", "cpp", "no_device.cpp"),
           os.path.join(CSRC, "build.cpp"), os.path.join(CSRC, "layout.cpp"), os.path.join(CSRC, "archive.cpp"),
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-lpthread", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def _run(driver, mode, paths):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")
    out = []
    for k in range(0, len(paths), 200):
        r = subprocess.run([driver, mode, *map(str, paths[k:k + 200])], capture_output=True, timeout=600,
                           env=env, encoding="utf-8", errors="replace")
        assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
        assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
        out += r.stdout.splitlines()
    assert len(out) == len(paths)
    return out


def _mutations(data: bytes, rng: random.Random, n: int, binary: bool = False):
    out = []
    lines = data.split(b"\n")
    for _ in range(n):
        kind = rng.randrange(6)
        if kind == 0:
            m = data[: rng.randrange(len(data) + 1)]
        elif kind == 1:
            b = bytearray(data)
            for _ in range(rng.randrange(1, 6)):
                b[rng.randrange(len(b))] = rng.randrange(256)
            m = bytes(b)
        elif kind == 2:
            big = rng.choice([0, 2**32, 2**63, 2**64 - 1, 99999999999999999999])
            digits = [i for i, c in enumerate(data) if 48 <= c <= 57]
            if binary or not digits:  # a huge count or size in place of an aligned word
                i = 4 + 8 * rng.randrange(max(1, (len(data) - 12) // 8))
                m = data[:i] + (big % 2**64).to_bytes(8, "little") + data[i + 8:]
            else:
                i = rng.choice(digits)
                m = data[:i] + str(big).encode() + data[i:]
        elif kind == 3:
            ls = list(lines)
            del ls[rng.randrange(len(ls))]
            m = b"\n".join(ls)
        elif kind == 4:
            ls = list(lines)
            i = rng.randrange(len(ls))
            ls.insert(i, ls[i])
            m = b"\n".join(ls)
        else:
            ls = list(lines)
            rng.shuffle(ls)
            m = b"\n".join(ls)
        out.append(m)
    return out


def test_layouts_and_build_plan(driver, tmp_path):
    from test_failure_paths import test_malformed_layouts_are_rejected as case_fn
    (mark,) = case_fn.pytestmark  # parametrize("text,msg", [...])
    texts = [TOP0 + HEAD + "0\t0\t2\n1\t2\t1\n2\t3\t1\n3\t4\t1\n", TOP0 + HEAD + "0\t0\t1\n" + REST]
    texts += [text for text, _ in mark.args[1]]
    assert len(texts) > 10
    rng = random.Random(11)
    blobs = [t.encode() for t in texts]
    for t in texts[:2]:
        blobs += _mutations(t.encode(), rng, 150)
    paths = []
    for i, b in enumerate(blobs):
        p = tmp_path / f"l{i}.txt"
        p.write_bytes(b)
        paths.append(p)
    out = _run(driver, "layout", paths)
    assert out[0].endswith(": ok") and out[1].endswith(": ok")
    assert sum(line.endswith(": ok") for line in out) < len(out)  # the malformed ones are rejected


def test_config_and_layout_files(driver, tmp_path):
    with open(os.path.join(ROOT, "tests", "golden", "config_header.json")) as f:
        header = json.load(f)["read_from_input"]
    good = header + TOP0 + HEAD + "0\t0\t1\n" + REST
    rng = random.Random(12)
    blobs = [good.encode()] + _mutations(good.encode(), rng, 300)
    paths = []
    for i, b in enumerate(blobs):
        p = tmp_path / f"f{i}.txt"
        p.write_bytes(b)
        paths.append(p)
    out = _run(driver, "files", paths)
    assert out[0].endswith(": ok")


def test_archives(driver, tmp_path):
    ibf = CL.ibf_record(70, 9, 2, np.arange(2 * 9, dtype=np.uint64), np.arange(128, dtype=np.uint64), True)
    ibf_b = CL.ibf_bytes(ibf)
    small = CL.ibf_record(3, 4, 2, np.zeros(4, dtype=np.uint64))
    hibf_b = CL.hibf_bytes(CL.HibfRecord(3, [small], [np.zeros(3, dtype=np.uint64)], [(0, 0)],
                                         [np.arange(3, dtype=np.uint64)]))
    rng = random.Random(13)
    for kind, data in (("ibf", ibf_b), ("hibf", hibf_b)):
        blobs = [data] + [data[:cut] for cut in range(0, len(data), 3)] + _mutations(data, rng, 200, True)
        paths = []
        for i, b in enumerate(blobs):
            p = tmp_path / f"{kind}{i}.bin"
            p.write_bytes(b)
            paths.append(p)
        out = _run(driver, kind, paths)
        # the well-formed archive parses up to the device (none here); every truncation is an IO error
        assert out[0].split(": ", 1)[1].startswith("11 ")
        assert all(line.split(": ", 1)[1].startswith("12 ") for line in out[1:len(data) // 3 + 1])
