"""Builds libhibf_gpu.so (HIP kernels + C-ABI + host runtime) for gfx950.

Run `python -m hibf_amd.build` from the repository root, or call build().
The library is built in-tree next to this file so that it travels with the
repository to the machine that runs it.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhibf_gpu.so")

HIP_SOURCES = ["ibf_kernels.hip", "count_wide.hip", "hibf_kernels.hip", "hibf_wave.hip", "serving.hip", "server.hip", "scan.hip",
               "kmer.hip", "hll.hip", "hll_union.hip", "capi.hip"]
CXX_SOURCES = ["layout.cpp", "build.cpp", "archive.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function",
          f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
# Packed fp32 instructions stay off in device code.
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in HIP_SOURCES + CXX_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(HERE, "_build", src + ".o")
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libhibf_gpu.so failed")
    os.replace(tmp, LIB)
    return LIB


CPP_TESTS = [os.path.join(ROOT, "tests", "cpp", "snippets.cpp"), os.path.join(ROOT, "tests", "cpp", "archive_roundtrip.cpp"),
             os.path.join(ROOT, "tests", "cpp", "hibf_members.cpp"), os.path.join(ROOT, "tests", "cpp", "hll_device.cpp"),
             os.path.join(ROOT, "tests", "cpp", "compute_layout.cpp")]


def build_cpp_tests(verbose: bool = False) -> list[str]:
    """Compiles the C++ programs that use the seqan::hibf headers (host g++,
    linked against libhibf_gpu.so)."""
    out = []
    for src in CPP_TESTS:
        exe = os.path.splitext(src)[0]
        cmd = ["g++", "-std=c++20", "-O1", "-Wall", f"-I{os.path.join(ROOT, 'include')}", src, f"-L{HERE}",
               "-lhibf_gpu", "-Wl,-rpath,$ORIGIN/../../hibf_amd", "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed on {src}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_tests(verbose=True))
