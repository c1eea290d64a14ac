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
from concurrent.futures import ThreadPoolExecutor

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


def _compile(cmd: list[str]) -> tuple[int, str]:
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass reports the device-only target feature as unknown
    noise = ("packed-fp32-ops",)
    msgs = [ln for ln in (res.stdout + res.stderr).splitlines()
            if ln.strip() and not any(n in ln for n in noise)]
    return res.returncode, "\n".join(msgs)


def build(verbose: bool = False, csrc: str = CSRC, out: str | None = None) -> str:
    """Compile the library (always recompiles) and return its path.

    Each translation unit compiles to an object in parallel (they share no
    device code), then one link makes the shared library. ``csrc`` and
    ``out`` name another source folder and output path (default: this
    package's)."""
    sources = sorted(glob.glob(os.path.join(csrc, "*.hip")))
    if not sources:
        raise RuntimeError(f"no HIP sources under {csrc}")
    out = out or library_path()
    os.makedirs(os.path.dirname(out), exist_ok=True)
    objdir = os.path.join(os.path.dirname(out), "obj")
    os.makedirs(objdir, exist_ok=True)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    inc = f"-I{os.path.join(os.path.dirname(HERE), 'include')}"
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objs, cmds = [], []
    for src in sources:
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        objs.append(obj)
        cmds.append([hipcc(), *compile_flags, inc, "-c", src, "-o", obj])
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(_compile, cmds))
    for cmd, (rc, msg) in zip(cmds, results):
        if verbose:
            print(" ".join(cmd), flush=True)
            if msg:
                print(msg)
        if rc != 0:
            raise RuntimeError("hipcc failed:\n" + msg)
    tmp = out + ".tmp"
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
            f"-Wl,-rpath,{rocm}/lib"]
    rc, msg = _compile(link)
    if verbose:
        print(" ".join(link), flush=True)
    if rc != 0:
        raise RuntimeError("link failed:\n" + msg)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
