"""Builds the gfx950 shared library overlap_amd/libovl.so in place with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libovl.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("ovl_kernels.hip", "ovl_api.hip")]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # keep hipcc from forming packed FP32 instructions
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    cmd = ["hipcc", *FLAGS, *SOURCES, "-o", LIB + ".tmp", "-lrocsolver", "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_compat_test(verbose: bool = False) -> str:
    """C++ test driver of the reference-shaped API (tests/compat), linked against libovl.so and
    the oracle; needs oracle/liboracle.so to exist."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "compat", "test_compat.cpp")
    out = os.path.join(root, "tests", "compat", "test_compat")
    cmd = ["g++", "-O2", "-std=c++17", src, "-o", out, f"-L{HERE}", f"-L{os.path.join(root, 'oracle')}",
           "-lovl", "-loracle", "-Wl,-rpath,$ORIGIN/../../overlap_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compat test build failed")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
