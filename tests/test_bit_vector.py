"""seqan::hibf::bit_vector (include/hibf/misc/bit_vector.hpp), the result type
of the IBF's bulk_contains, on the host under AddressSanitizer and UBSan:
the expectations of the reference's test/unit/hibf/bit_vector_test.cpp, the
iterator concepts, and random operation sequences against a word-level model
(tests/cpp/bit_vector_check.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GROUPS = ["construction_and_assignment", "iterators_and_access", "all_any_none", "modifiers", "word_wise_operators",
          "flips", "size_empty_output", "data_layout", "random_against_model"]


def test_bit_vector(tmp_path):
    exe = tmp_path / "bit_vector_check"
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-Wall", "-Wextra", "-Werror", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", f"-I{ROOT}/include",
                        os.path.join(ROOT, "tests", "cpp", "bit_vector_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == [f"{g} ok" for g in GROUPS]
