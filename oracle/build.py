"""ORACLE build recipe (test infrastructure only).

Compiles the CPU FP64 restatement oracle/icsf_oracle.cpp with g++ into
oracle/liboracle.so.  -ffp-contract=off and no -march flags: no fused
multiply-adds, so every expression evaluates exactly as written.
The reference implementation itself (C++ on OpenFOAM-v2112) cannot be built
here: OpenFOAM is not installed, so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


FAST_LIB = os.path.join(HERE, "liboracle_fast.so")
FAST_FLAGS = ["-O3", "-march=x86-64-v3"]


def build(openmp: bool = False, fast: bool = False) -> str:
    """fast: the timing build of the CPU baseline (-O3, AVX2/FMA, contraction
    allowed; x86-64-v3 rather than -march=native so that a library built on
    one host runs on another); otherwise the bit-exact checker build."""
    src = os.path.join(HERE, "icsf_oracle.cpp")
    if fast:
        out = FAST_LIB
        cmd = ["g++", *FAST_FLAGS, "-std=c++17", "-fPIC", "-shared", src, "-o", out + ".tmp"]
    else:
        out = LIB if not openmp else os.path.join(HERE, "liboracle_omp.so")
        cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", src, "-o", out + ".tmp"]
    if openmp:
        cmd.insert(1, "-fopenmp")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
