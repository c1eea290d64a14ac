"""Build the HIP library libmaddening_lbm.so in-tree (gfx950).

``python -m maddening_amd.build`` compiles every ``csrc/*.hip`` into
``maddening_amd/lib/libmaddening_lbm.so``. hipcc cross-compiles gfx950
without a GPU, so this runs anywhere ROCm is installed.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libmaddening_lbm.so"
ARCH = os.environ.get("MADDENING_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 VALU forms stay off: plain scalar FP32 code throughout
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # no implicit a*b+c fusion: every kernel variant (single lattice, row,
    # batched) evaluates the collision with the same roundings
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def library_path() -> str:
    return os.path.join(LIBDIR, LIBNAME)


def build(verbose: bool = False) -> str:
    """Compile the library (always recompiles) and return its path."""
    os.makedirs(LIBDIR, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    out = library_path()
    tmp = out + ".tmp"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{os.path.join(os.path.dirname(HERE), 'include')}",
           *sources, "-o", tmp, f"-Wl,-rpath,{rocm}/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass reports the device-only target feature as unknown
    noise = ("packed-fp32-ops",)
    msgs = [ln for ln in (res.stdout + res.stderr).splitlines()
            if ln.strip() and not any(n in ln for n in noise)]
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
