"""Build recipe for the C restatement (oracle/lbm_ref.c) -> oracle/lib/liblbm_oracle.so.

Test infrastructure: the library is loaded only by tests/ and bench.py's
CPU-baseline leg. ``python -m oracle.build`` compiles it with gcc and
OpenMP for any x86-64 host with AVX2 (the build machine and the machine
that runs the GPU need not be the same).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lbm_ref.c")
LIB = os.path.join(HERE, "lib", "liblbm_oracle.so")
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
         "-Wall", "-Wextra"]


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), *FLAGS, "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("gcc failed:\n" + res.stdout + res.stderr)
    if verbose and (res.stdout or res.stderr):
        print(res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
