"""The order in which split user bins are chunked, without a GPU.

The reference fills a split user bin's technical bins with chunks of its
k-mer set's iteration order (src/build/insert_into_ibf.cpp:28-47), a
robin_hood::unordered_flat_set<uint64_t> whose order follows from everything
the set held before (clear() keeps its table).  Checked here:
- the oracle's restatement of that set (orc_fs) and the library's
  (hibf_amd/csrc/flat_set_order.hpp) print the golden orders the reference's
  own set printed (tests/golden/robin_hood_orders.json), and, where the
  reference tree is present, the orders it prints for random scripts;
- the library's build plan, which replays only what can change a split bin's
  order, gives every split user bin the chunks that a full replay of the
  reference's hierarchical_build on the oracle's set gives
  (tests/hibf_helpers.reference_split_chunks), on layouts with splits on
  every level, after merged fullest bins, and below them."""
import json
import os
import random
import subprocess

import numpy as np
import pytest

import oracle as O
from hibf_amd import native
from hibf_helpers import reference_split_chunks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "robin_hood_orders.json")
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "robin_hood_ref")
CSRC = os.path.join(ROOT, "hibf_amd", "csrc")


def _compile(tmp_path_factory, name, sources):
    exe = tmp_path_factory.mktemp(name) / name
    r = subprocess.run(["g++", "-std=c++20", "-O2", "-Wall", f"-I{CSRC}", f"-I{ROOT}/include", *sources, "-o",
                        str(exe), "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


@pytest.fixture(scope="module")
def library_set_script(tmp_path_factory):
    return _compile(tmp_path_factory, "flat_set_script", [os.path.join(ROOT, "tests", "cpp", "flat_set_script.cpp")])


@pytest.fixture(scope="module")
def library_plan(tmp_path_factory):
    return _compile(tmp_path_factory, "split_plan", [os.path.join(ROOT, "tests", "cpp", "split_plan.cpp"),
                                                     os.path.join(CSRC, "build.cpp"),
                                                     os.path.join(CSRC, "layout.cpp")])


def _run(exe, script):
    r = subprocess.run([exe], input=script, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _goldens():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _goldens(), ids=lambda c: c["name"])
def test_oracle_set_prints_the_reference_orders(case):
    assert O.flat_set_script(case["script"]) == case["output"]


@pytest.mark.parametrize("case", _goldens(), ids=lambda c: c["name"])
def test_library_set_prints_the_reference_orders(case, library_set_script):
    assert _run(library_set_script, case["script"]) == case["output"]


def _random_script(rng, n_ops):
    lines = []
    for _ in range(n_ops):
        s = rng.randrange(0, 6)
        r = rng.random()
        if r < 0.45:
            lines.append(f"gen {s} {rng.randrange(1, 1000)} {rng.randrange(0, 4000)} 0 "
                         f"{rng.choice([1, 2, 5, 7, 9, 50, 300, 3000, 14304, 70000, 260000])}")
        elif r < 0.6:
            lines.append(f"clear {s}")
        elif r < 0.75:
            b = rng.randrange(0, 6)
            if b != s:
                lines.append(f"merge {s} {b}")
        elif r < 0.9:
            lines.append(f"genmod {s} {rng.randrange(1, 9)} {rng.randrange(1, 50)} {rng.choice([100, 5000, 90000])} "
                         f"{rng.choice([1, 10, 777, 40000])}")
        else:
            lines.append(f"vals {s} " + " ".join(str(rng.getrandbits(64)) for _ in range(rng.randrange(1, 40))))
        lines.append(f"dump {s} {int(rng.random() < 0.1)}")
    return "\n".join(lines) + "\n"


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="oracle/_ref/robin_hood_ref is built only where the "
                                                        "reference tree is present (make -C oracle ref)")
@pytest.mark.parametrize("seed", range(6))
def test_random_scripts_against_the_reference_set(seed, library_set_script):
    script = _random_script(random.Random(seed), 60)
    want = _run(REF_EXE, script)
    assert O.flat_set_script(script) == want
    assert _run(library_set_script, script) == want


# ------------------------------------------------------------- the plan ----

def _layout(sizes, tmax):
    c = native.validate_config(native.make_config(number_of_user_bins=len(sizes), tmax=tmax))
    return native.compute_layout(c, sizes)


def _plan_chunks(exe, tmp_path, layout, sizes, seed, tmax, threads=3):
    lp, sp = tmp_path / "layout.txt", tmp_path / "sizes.txt"
    lp.write_text(layout)
    sp.write_text("\n".join(str(int(s)) for s in sizes) + "\n")
    r = subprocess.run([exe, str(lp), str(sp), str(seed), str(threads), str(tmax), "full"], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        f = line.split()
        if f[0] == "seconds":
            continue
        out[int(f[0])] = np.array([int(x) for x in f[4:]], dtype=np.int64)
    return out


def _full_replay(layout, sizes, seed):
    return reference_split_chunks(layout, lambda s, u: s.insert_synth(seed, u, 0, int(sizes[u])),
                                  lambda u, keys: _synth_index(seed, u, sizes[u], keys))


def _synth_index(seed, u, n, keys):
    from hibf_amd.synth import synth_values
    vals = synth_values(seed, u, np.arange(int(n), dtype=np.uint64))
    # keys are the set's iteration order of exactly these values: match the two sorted orders
    assert keys.size == vals.size
    out = np.empty(keys.size, dtype=np.int64)
    out[np.argsort(keys)] = np.argsort(vals)
    return out


# Handwritten three-level layout: the root's fullest bin is merged and split user bins follow it; IBF [0]'s
# fullest bin is merged again (IBF [0;0]) and split bins follow that too, so the root's order depends on the
# whole subtree below [0]; IBF [1] keeps a split fullest bin and more split bins.
_DEEP = """#TOP_LEVEL_IBF fullest_technical_bin_idx:0
#LOWER_LEVEL_IBF_0 fullest_technical_bin_idx:0
#LOWER_LEVEL_IBF_1 fullest_technical_bin_idx:0
#LOWER_LEVEL_IBF_0;0 fullest_technical_bin_idx:2
#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS
0\t0;0;0\t1;1;2
1\t0;0;2\t1;1;3
2\t0;0;5\t1;1;1
3\t0;1\t1;3
4\t0;4\t1;2
5\t0;6\t1;1
6\t1;0\t1;4
7\t1;4\t1;2
8\t1;6\t1;1
9\t2\t1
10\t3\t3
11\t6\t2
12\t8\t1
13\t9\t4
"""


def test_plan_matches_full_replay_on_a_deep_handwritten_layout(library_plan, tmp_path):
    sizes = np.array([900, 4000, 300, 2500, 1800, 700, 5000, 1200, 400, 3000, 9000, 2200, 100, 6000],
                     dtype=np.uint64)
    got = _plan_chunks(library_plan, tmp_path, _DEEP, sizes, 21, 64)
    want = _full_replay(_DEEP, sizes, 21)
    assert sorted(got) == sorted(want) == [0, 1, 3, 4, 6, 7, 10, 11, 13]
    for u in want:
        assert np.array_equal(got[u], want[u]), u


@pytest.mark.parametrize("n_ub,tmax,median,sigma,seed", [
    (200, 16, 2000, 1.5, 1),
    (700, 64, 500, 1.2, 2),
    (1500, 64, 300, 2.0, 3),
    (3000, 64, 200, 1.5, 4),
    (400, 8, 1000, 1.0, 5),
])
def test_plan_matches_full_replay_on_computed_layouts(library_plan, tmp_path, n_ub, tmax, median, sigma, seed):
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, size=n_ub)), 5, 200_000).astype(np.uint64)
    layout = _layout(sizes, tmax)
    got = _plan_chunks(library_plan, tmp_path, layout, sizes, 100 + seed, tmax)
    want = _full_replay(layout, sizes, 100 + seed)
    assert len(want) > 0 and sorted(got) == sorted(want)
    for u in want:
        assert np.array_equal(got[u], want[u]), u
    # the split bins' orders are not their value ranges (the old chunking)
    assert any(not np.array_equal(want[u], np.sort(want[u])) for u in want)
