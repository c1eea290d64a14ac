"""Throughput of the D3Q19 lattice-Boltzmann step on one MI355X.

Workload: D3Q19 BGK, periodic 256^3 Taylor-Green vortex (u0 = 0.05,
tau = 0.8), float32 populations. One "step" is one full lattice update
(collision + streaming over all 16,777,216 cells). The timed region is ONE
call of lbm_step(nsteps=K) with the state already resident in HBM, moments
of the final state included.

Prints one JSON line. `value` is MLUPS (million lattice-cell updates per
second). `roofline` prices the dominant kernel at the algorithmic
2 * Q * 4 = 152 bytes per cell and launch against 8 TB/s HBM3E;
`cpu_baseline` times the float32 NumPy restatement of the reference step on
this host (the oracle, used as a port of the reference's CPU path).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))


def taylor_green_device(n, u0, device):
    import torch
    k = 2.0 * np.pi / n
    x = torch.arange(n, device=device, dtype=torch.float64) * k
    X = x.view(n, 1, 1)
    Y = x.view(1, n, 1)
    Z = x.view(1, 1, n)
    u = torch.empty((n, n, n, 3), device=device, dtype=torch.float32)
    u[..., 0] = (u0 * torch.sin(X) * torch.cos(Y) * torch.cos(Z)).float()
    u[..., 1] = (-u0 * torch.cos(X) * torch.sin(Y) * torch.cos(Z)).float()
    u[..., 2] = 0.0
    dp = (u0 ** 2 / 16.0) * (torch.cos(2 * X) + torch.cos(2 * Y)) * (torch.cos(2 * Z) + 2.0)
    rho = (1.0 + 3.0 * dp).float()
    return rho, u


def time_kernel(kern, f0, steps, warmup):
    """Returns (seconds for `steps` steps, launches in the timed call)."""
    import torch
    out = kern.empty()
    work = kern.workspace(max(steps, warmup, 2))
    if warmup:
        kern.step(f0, warmup, out=out, work=work, moments=False)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    kern.step(f0, steps, out=out, work=work, moments=True)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = start.elapsed_time(end) / 1e3
    launches = steps + 1  # pull: collide + (K-1) fused + stream; push: K + moments pass
    return ev, wall, launches, out


def cpu_baseline(target_seconds=15.0):
    """Float32 NumPy restatement of the reference step (the oracle) on a
    64^3 Taylor-Green sample; single-threaded NumPy."""
    sys.path.insert(0, ROOT)
    from oracle import lbm_numpy as orc
    from maddening_amd import workloads
    n = 64
    rho, u = workloads.taylor_green_3d(n, 0.05)
    e, w, _ = orc.lattice(19)
    f = orc.equilibrium(rho, u, e, w).astype(np.float32)
    orc.update(f, tau=0.8)  # warm-up
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < target_seconds:
        f = orc.update(f, tau=0.8)["f"]
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": round(n ** 3 * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": 1,
            "kind": "port",
            "sample": f"D3Q19 64^3 Taylor-Green, float32 NumPy restatement of lbm.py:899-986 "
                      f"(AoS, roll streaming), {steps} steps in {dt:.1f} s"}


def read_traffic():
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", "pmc_d3q19_256_f32.json")
    try:
        with open(path) as fh:
            return json.load(fh).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--scheme", default="auto")
    ap.add_argument("--block-x", type=int, default=0)
    ap.add_argument("--row-vec", type=int, default=0)
    ap.add_argument("--xcd-remap", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB elementwise copy (torch.mul by 1) so HBM "
                         "byte counters can be scaled by a kernel of known traffic")
    ap.add_argument("--variants", default="f16dev,bf16dev",
                    help="extra storage types to time (comma list, '' for none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is out of scope for this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    from maddening_amd.lbm import LBMKernel

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = args.size
    cells = n ** 3
    rho, u = taylor_green_device(n, 0.05, dev)

    def make(storage):
        return LBMKernel((n, n, n), lattice="D3Q19", tau=0.8, storage=storage,
                         scheme=args.scheme, block_x=args.block_x, row_vec=args.row_vec,
                         xcd_remap=args.xcd_remap, device=dev)

    if args.calibrate_copy:
        a = torch.ones(1 << 28, device=dev)  # 1 GiB float32
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b

    kern = make("f32")
    f0 = kern.equilibrium(rho, u)
    ev, wall, launches, out = time_kernel(kern, f0, args.steps, args.warmup)
    mlups = cells * args.steps / ev / 1e6
    b_cell = 2 * 19 * 4
    per_launch_s = ev / launches
    achieved = b_cell * cells / per_launch_s / 1e9
    result = {
        "metric": "D3Q19 MLUPS & %HBM peak (fp32 vs fp16-dev storage); fwd+adjoint ms/step",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ev / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"D3Q19 BGK periodic {n}^3 Taylor-Green vortex, u0=0.05, "
                               f"tau=0.8, float32 populations",
                   "cells": cells, "scheme": kern.scheme, "wall_s_timed_call": round(wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(),
                     "kernel": "k_cell<D3Q19,f32,pull+collide> (152 B/cell algorithmic)",
                     "avg_launch_us": round(per_launch_s * 1e6, 2)},
    }
    extra = {}
    for storage in [s for s in args.variants.split(",") if s]:
        k2 = make(storage)
        f2 = k2.equilibrium(rho, u)
        ev2, _, launches2, _ = time_kernel(k2, f2, args.steps, args.warmup)
        bc2 = 2 * 19 * (2 if storage.endswith("dev") else 4)
        extra[storage] = {"MLUPS": round(cells * args.steps / ev2 / 1e6, 1),
                          "GB/s": round(bc2 * cells / (ev2 / launches2) / 1e9, 1),
                          "speedup_vs_f32": round(ev / ev2, 3)}
        del f2, k2
    if extra:
        result["variants"] = extra
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
