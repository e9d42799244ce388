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
                "hibf/misc/print.hpp", "hibf/misc/timer.hpp", "hibf_gpu.h"):
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
    timers and copy semantics of the C++ HIBF class."""
    exe = os.path.join(ROOT, "tests", "cpp", "hibf_members")
    assert os.path.exists(exe), "build tests/cpp/hibf_members first (python -m hibf_amd.build)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == 7 and all(line.endswith(" ok") for line in lines), r.stdout
