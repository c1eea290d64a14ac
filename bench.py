"""Throughput of the D3Q19 lattice-Boltzmann step on one MI355X.

Workload: D3Q19 BGK, periodic 256^3 Taylor-Green vortex (u0 = 0.05,
tau = 0.8), float32 populations stored as f - w (same bytes as plain f32,
whose rate is reported beside it). One "step" is one full lattice update
(collision + streaming over all 16,777,216 cells). The timed region is ONE
call of lbm_step(nsteps=K) with the state already resident in HBM, moments
of the final state included.

Prints one JSON line. `value` is MLUPS (million lattice-cell updates per
second). `roofline` prices the dominant kernel at the algorithmic
2 * Q * 4 = 152 bytes per cell update, times the updates one launch makes
(two for the two-step launch), against 8 TB/s HBM3E, with the launch's
measured HBM traffic and rate beside it; `cpu_baseline` times the C
restatement of the reference step (oracle/lbm_ref.c, OpenMP) on this
host's cores, with the float32 NumPy restatement on one core beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))
STORAGE_BYTES = {"f32": 4, "f32dev": 4, "f64": 8, "f16dev": 2, "bf16dev": 2}


def fused_kernel_label(storage, row_vec, pairs=0):
    """Name of the kernel that runs the fused pull+collide launches."""
    if pairs:
        return f"k_pair<D3Q19,{storage}> (two steps per launch)"
    if row_vec > 0:
        return f"k_row<D3Q19,{storage},V={row_vec}>"
    return f"k_cell<D3Q19,{storage},pull+collide>"


def auto_row_vec(storage, nc, row_vec):
    """Row width the library picks for a periodic pull run (lbm_forward.hip
    row_width): 2 cells per lane for 16-bit storage when rows split into
    128-cell wave segments, the per-cell kernel otherwise."""
    if row_vec < 0:
        return 0
    if row_vec == 0:
        if STORAGE_BYTES[storage] != 2:
            return 0
        return 2 if nc % 128 == 0 else (1 if nc % 64 == 0 else 0)
    return row_vec if nc % (64 * row_vec) == 0 else 0


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


def load_kernels(kern, steps):
    """Run the timed call's launch sequence once on an 8-cell-thick copy of
    the lattice, so every kernel it uses is loaded on the device before the
    timed region (a K-step call can use launches a shorter warm-up does not).
    The copy is a different, tiny lattice: no workload step runs here."""
    import torch
    from maddening_amd.lbm import LBMKernel
    shape = (8,) + tuple(kern.grid_shape[1:]) if kern.D == 3 else kern.grid_shape
    k = LBMKernel(shape, lattice=kern.lat.name, tau=kern.tau, storage=kern.storage,
                  scheme=kern.scheme, block_x=kern.block_x, row_vec=kern.row_vec,
                  xcd_remap=kern.xcd_remap, pair="force" if kern.pair != "off" else "off",
                  tri=kern.tri, device=kern.device)
    k.step(k.equilibrium(), steps)
    torch.cuda.synchronize()


def time_kernel(kern, f0, steps, warmup):
    """Returns (seconds for `steps` steps, launches in the timed call).

    The warm-up writes the moments too, so the timed call's rho/u/p buffers
    come from torch's allocator cache instead of a fresh device allocation."""
    import torch
    out = kern.empty()
    work = kern.workspace(max(steps, warmup, 2))
    load_kernels(kern, steps)
    if warmup:
        kern.step(f0, warmup, out=out, work=work, moments=True)
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
    # pull: collide, two-step launches, single fused launches, stream + moments
    launches, pairs = kern.schedule(steps)
    return ev, wall, launches, pairs, out


def time_pair_launch(kern, f0, reps=40):
    """Average duration of one two-step launch (lbm_pull_phase phase 3, the
    kernel lbm_step runs for pairs of steps), timed with HIP events on the
    stream the library launches on (torch's current stream)."""
    import ctypes
    import torch
    from maddening_amd import _lib
    from maddening_amd.lbm import _stream
    d = kern.desc()
    L = _lib.lib()
    s = _stream(kern.device)
    a, b = f0.clone(), kern.empty()

    def launch(src, dst):
        _lib.check(L.lbm_pull_phase(s, ctypes.byref(d), 3, ctypes.c_void_p(src.data_ptr()),
                                    ctypes.c_void_p(dst.data_ptr()), None, None, None, None,
                                    None, None, None), "lbm_pull_phase")

    for i in range(4):
        launch(a, b) if i % 2 == 0 else launch(b, a)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for i in range(reps):
        launch(a, b) if i % 2 == 0 else launch(b, a)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3 / reps


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(target_seconds=15.0, n=256, numpy_seconds=5.0):
    """The reference step timed on the host cores, beside the GPU run.

    Main figure: the C restatement of lbm.py:899-986 (oracle/lbm_ref.c, AoS
    (*G, Q) float32 like the reference's state, OpenMP over cells) on the
    same 256^3 Taylor-Green workload, for about `target_seconds` of steps.
    Threads: the cores this process may run on, capped by OMP_NUM_THREADS
    when that is set. Also reported: the float32 NumPy restatement, which
    keeps the reference's array-op structure (roll streaming, broadcast Guo
    source), on one core at 64^3."""
    sys.path.insert(0, ROOT)
    from oracle import lbm_c as occ
    from oracle import lbm_numpy as orc
    from maddening_amd import workloads
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    e, w, _ = orc.lattice(19)
    rho, u = workloads.taylor_green_3d(n, 0.05)
    f = orc.equilibrium(rho.astype(np.float32), u.astype(np.float32), e, w.astype(np.float32))
    f = np.ascontiguousarray(f, dtype=np.float32)
    del rho, u
    # buffers touched before timing, so page faults stay out of the sample
    bufs = (f.copy(), np.zeros_like(f), np.zeros_like(f))
    occ.run(f, 1, tau=0.8, nthreads=threads, buffers=bufs)      # warm-up
    t0 = time.perf_counter()
    occ.run(f, 2, tau=0.8, nthreads=threads, buffers=bufs)      # per-step estimate
    per_step = (time.perf_counter() - t0) / 2
    steps = max(2, int(target_seconds / max(per_step, 1e-3)))
    t0 = time.perf_counter()
    occ.run(f, steps, tau=0.8, nthreads=threads, buffers=bufs)
    dt = time.perf_counter() - t0
    del bufs
    res = {"value": round(n ** 3 * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads,
           "kind": "port",
           "sample": f"D3Q19 {n}^3 Taylor-Green (u0=0.05, tau=0.8), float32, C restatement of "
                     f"lbm.py:899-986 (AoS, OpenMP over cells), {steps} steps in {dt:.1f} s",
           "cpu_model": _cpu_model(), "affinity_cores": affinity, "omp_num_threads": omp}
    del f
    m = 64
    rho, u = workloads.taylor_green_3d(m, 0.05)
    g = orc.equilibrium(rho, u, e, w).astype(np.float32)
    orc.update(g, tau=0.8)
    nsteps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < numpy_seconds:
        g = orc.update(g, tau=0.8)["f"]
        nsteps += 1
    dt = time.perf_counter() - t0
    res["numpy_1_core"] = {
        "value": round(m ** 3 * nsteps / dt / 1e6, 3), "unit": "MLUPS", "cores": 1,
        "sample": f"D3Q19 {m}^3 Taylor-Green, float32 NumPy restatement (roll streaming), "
                  f"{nsteps} steps in {dt:.1f} s"}
    return res


def precision_leg(n, dev, checkpoints=(100, 1000, 5000)):
    """Deviation storage vs the float64 kernel on the 256^3 Taylor-Green case:
    max |du| / u0 and max |drho| at each checkpoint, plus the float64 sums
    of mass and momentum (drift of total mass relative to the start)."""
    import torch
    from maddening_amd.lbm import LBMKernel
    u0 = 0.05
    rho, u = taylor_green_device(n, u0, dev)
    ref = {}
    out = {}
    for storage in ("f64", "f32", "f32dev", "f16dev", "bf16dev"):
        k = LBMKernel((n, n, n), tau=0.8, storage=storage, device=dev)
        f = k.equilibrium(rho, u)
        m0 = k.conservation(f)
        work = k.workspace(max(checkpoints))
        done = 0
        rows = {}
        for c in checkpoints:
            f, r, v, _ = k.step(f, c - done, work=work)
            done = c
            m = k.conservation(f)
            if storage == "f64":
                ref[c] = (r, v)
                rows[c] = {"mass_drift": float((m[0] - m0[0]) / m0[0])}
            else:
                rr, vr = ref[c]
                rows[c] = {"max_du_over_u0": float((v - vr).abs().max() / u0),
                           "max_drho": float((r - rr).abs().max()),
                           "rel_du": float((v - vr).abs().max() / vr.abs().max()),
                           "mass_drift": float((m[0] - m0[0]) / m0[0])}
        out[storage] = rows
        del f, work, k
        torch.cuda.empty_cache()
    return out


def adjoint_leg(dev, nsteps=500, shape=(128, 128, 256), fd_cells=16, reps=3):
    """Sphere-packed channel (porosity 0.7, radii 6-12, seed 0), tau field 0.7,
    body force (0, 0, 1e-5), loss = mean u_z over the outlet plane. Times the
    float32 forward and the sqrt(N)-checkpointed adjoint, and (f32) the
    adjoint with every state kept, which is what the rollout's default
    ("auto") does when the states fit in free device memory, as here; checks
    the float64 adjoint against central differences of the float64 forward."""
    import torch
    from maddening_amd import workloads
    from maddening_amd.adjoint import rollout
    from maddening_amd.lbm import LBMKernel
    mask = torch.from_numpy(workloads.sphere_packing(shape, porosity=0.7, rmin=6, rmax=12,
                                                     seed=0)).to(dev)
    F0 = 1e-5
    res = {"workload": f"D3Q19 {shape} sphere-packed channel, porosity "
                       f"{float((mask == 0).float().mean()):.3f}, tau field 0.7, F=(0,0,{F0}), "
                       f"{nsteps} steps, loss = mean u_z over the outlet plane"}

    def uz_loss(k, f, Fz):
        e = torch.as_tensor(k.lat.e, dtype=torch.float64, device=dev)
        fl = f[..., -1].double()
        rh = fl.sum(0)
        uz = ((fl * e[:, 2].view(-1, 1, 1)).sum(0) + 0.5 * Fz) / rh
        return torch.where(mask[..., -1] != 0, torch.zeros_like(uz), uz).mean()

    for storage in ("f32", "f64"):
        k = LBMKernel(shape, tau=0.7, storage=storage, bb="reference", device=dev)
        f0 = k.equilibrium()
        tau = torch.full(shape, 0.7, device=dev)
        # forward alone and forward + adjoint through the autograd rollout,
        # interleaved `reps` times (medians reported): clocks drift by several
        # percent between runs on a loaded box. One untimed rollout first so
        # the checkpoint and segment buffers come from torch's cache.
        work = k.workspace(nsteps)
        k.step(f0, 10, mask=mask, force=[0, 0, F0], tau_field=tau, work=work)
        Fw = torch.tensor([0.0, 0.0, F0], dtype=torch.float64, device=dev, requires_grad=True)
        tw = tau.clone().requires_grad_(True)
        fw, _, _, _ = rollout(k, f0, nsteps, force=Fw, tau_field=tw, mask=mask,
                              checkpoint="sqrt")
        uz_loss(k, fw, Fw[2]).backward()
        del fw, Fw, tw
        torch.cuda.synchronize()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        fwd_t, roll_t, adj_t, tot_t = [], [], [], []
        for _ in range(reps):
            e0.record()
            k.step(f0, nsteps, mask=mask, force=[0, 0, F0], tau_field=tau, work=work)
            e1.record()
            torch.cuda.synchronize()
            fwd_t.append(e0.elapsed_time(e1) / nsteps)
            Fg = torch.tensor([0.0, 0.0, F0], dtype=torch.float64, device=dev,
                              requires_grad=True)
            tg = tau.clone().requires_grad_(True)
            torch.cuda.synchronize()
            e0.record()
            fN, _, _, _ = rollout(k, f0, nsteps, force=Fg, tau_field=tg, mask=mask,
                                  checkpoint="sqrt")
            e1.record()
            L = uz_loss(k, fN, Fg[2])
            L.backward()
            e2.record()
            torch.cuda.synchronize()
            roll_t.append(e0.elapsed_time(e1) / nsteps)
            adj_t.append(e1.elapsed_time(e2) / nsteps)
            tot_t.append(e0.elapsed_time(e2) / nsteps)
        fwd_ms, adj_ms = statistics.median(fwd_t), statistics.median(adj_t)
        rec = {"checkpoint": "sqrt",
               "forward_ms_per_step": round(fwd_ms, 4),
               "rollout_forward_ms_per_step": round(statistics.median(roll_t), 4),
               "adjoint_ms_per_step": round(adj_ms, 4),
               "adjoint_over_forward": round(adj_ms / fwd_ms, 3),
               "fwd_plus_adjoint_ms_per_step": round(statistics.median(tot_t), 4),
               "reps": reps,
               "forward_ms_range": [round(min(fwd_t), 4), round(max(fwd_t), 4)],
               "adjoint_ms_range": [round(min(adj_t), 4), round(max(adj_t), 4)],
               "loss": float(L), "dL_dF": float(Fg.grad[2])}
        if storage == "f32":
            # every post-collision state kept (N x 319 MB): no recompute
            del fN
            torch.cuda.empty_cache()
            from maddening_amd.adjoint import resolve_checkpoint
            # what the rollout's default ("auto") picks for this call here
            rec["default_checkpoint"] = resolve_checkpoint(
                k.ncells * 19 * 4, nsteps, "auto", device=dev)
            # one untimed all-states rollout first, so the timed one takes its
            # state buffers from torch's cache instead of fresh allocations
            for timed in (False, True):
                Fa = torch.tensor([0.0, 0.0, F0], dtype=torch.float64, device=dev,
                                  requires_grad=True)
                ta = tau.clone().requires_grad_(True)
                if timed:
                    break
                fA, _, _, _ = rollout(k, f0, nsteps, force=Fa, tau_field=ta, mask=mask,
                                      checkpoint="all")
                uz_loss(k, fA, Fa[2]).backward()
                del fA
            torch.cuda.synchronize()
            e0.record()
            fA, _, _, _ = rollout(k, f0, nsteps, force=Fa, tau_field=ta, mask=mask,
                                  checkpoint="all")
            e1.record()
            uz_loss(k, fA, Fa[2]).backward()
            e2.record()
            torch.cuda.synchronize()
            rec["all_states"] = {
                "adjoint_ms_per_step": round(e1.elapsed_time(e2) / nsteps, 4),
                "adjoint_over_forward": round(e1.elapsed_time(e2) / nsteps / fwd_ms, 3),
                "fwd_plus_adjoint_ms_per_step": round(e0.elapsed_time(e2) / nsteps, 4),
                "states_GB": round(nsteps * k.ncells * 19 * 4 / 1e9, 1),
                "dL_dF": float(Fa.grad[2]),
                "max_rel_diff_dL_dtau_vs_sqrt": float((ta.grad - tg.grad).abs().max()
                                                      / tg.grad.abs().max())}
            del fA, Fa, ta
            fN = None
            torch.cuda.empty_cache()
        if storage == "f64":
            def loss_at(Fz, tf):
                f, _, _, _ = k.step(f0, nsteps, mask=mask, force=[0, 0, Fz], tau_field=tf,
                                    moments=False, work=work)
                return float(uz_loss(k, f, Fz))
            h = 1e-4 * F0
            fd = (loss_at(F0 + h, tau) - loss_at(F0 - h, tau)) / (2 * h)
            rec["fd_dL_dF"] = fd
            rec["rel_err_dL_dF"] = abs(float(Fg.grad[2]) - fd) / abs(fd)
            errs = []
            fluid = (mask == 0).nonzero()
            g = torch.Generator().manual_seed(0)
            pick = torch.randperm(len(fluid), generator=g)[:fd_cells]
            gmax = float(tg.grad.abs().max())
            for i in pick.tolist():
                c = tuple(int(v) for v in fluid[i])
                tp, tm = tau.clone(), tau.clone()
                tp[c] += 1e-4
                tm[c] -= 1e-4
                hs = float(tp[c]) - float(tm[c])
                fdt = (loss_at(F0, tp) - loss_at(F0, tm)) / hs
                errs.append(abs(float(tg.grad[c]) - fdt) / max(abs(fdt), 1e-3 * gmax))
            rec["max_rel_err_dL_dtau"] = max(errs)
            rec["fd_cells"] = fd_cells
        res[storage] = rec
        del k, f0, work, fN
        torch.cuda.empty_cache()
    # the same channel under halfway bounce-back (the rollout reads link
    # words), float32, timing only
    k = LBMKernel(shape, tau=0.7, storage="f32", bb="halfway", device=dev)
    f0 = k.equilibrium()
    work = k.workspace(nsteps)
    k.step(f0, 10, mask=mask, force=[0, 0, F0], tau_field=tau, work=work)
    fwd_t, adj_t = [], []
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    for r in range(reps + 1):
        e0.record()
        k.step(f0, nsteps, mask=mask, force=[0, 0, F0], tau_field=tau, work=work)
        e1.record()
        torch.cuda.synchronize()
        Fh = torch.tensor([0.0, 0.0, F0], dtype=torch.float64, device=dev, requires_grad=True)
        th = tau.clone().requires_grad_(True)
        fH, _, _, _ = rollout(k, f0, nsteps, force=Fh, tau_field=th, mask=mask)
        L = uz_loss(k, fH, Fh[2])
        e2.record()
        torch.cuda.synchronize()
        e0b = torch.cuda.Event(enable_timing=True)
        e0b.record()
        L.backward()
        e2.record()
        torch.cuda.synchronize()
        if r:  # the first round warms the allocator
            fwd_t.append(e0.elapsed_time(e1) / nsteps)
            adj_t.append(e0b.elapsed_time(e2) / nsteps)
        del fH, Fh, th, L
    fwd_ms, adj_ms = statistics.median(fwd_t), statistics.median(adj_t)
    res["halfway_f32"] = {"forward_ms_per_step": round(fwd_ms, 4),
                          "adjoint_ms_per_step": round(adj_ms, 4),
                          "adjoint_over_forward": round(adj_ms / fwd_ms, 3), "reps": reps}
    del k, f0, work
    torch.cuda.empty_cache()
    return res


def time_phase(kern, phase, src, reps=20, mask=None, force_field=None, tau_field=None,
               moments=False):
    """Average duration of one lbm_pull_phase launch of `phase`, HIP events
    on the stream the library launches on (torch's current stream)."""
    import ctypes
    import torch
    from maddening_amd import _lib
    from maddening_amd.lbm import _stream
    d = kern.desc()
    L = _lib.lib()
    s = _stream(kern.device)
    dst = kern.empty()
    mom = [None, None, None]
    if moments:
        mom = [torch.empty(kern.grid_shape, device=kern.device),
               torch.empty(kern.grid_shape + (kern.D,), device=kern.device),
               torch.empty(kern.grid_shape, device=kern.device)]

    def p(t):
        return None if t is None else ctypes.c_void_p(t.data_ptr())

    def launch():
        _lib.check(L.lbm_pull_phase(s, ctypes.byref(d), phase, p(src), p(dst), p(mask), None,
                                    p(force_field), p(tau_field), *[p(t) for t in mom]),
                   "lbm_pull_phase")
    for _ in range(3):
        launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def node_update_leg(dev, nsteps=50, reps=3):
    """The drop-in call pattern: GraphManager.step, which calls
    LBMNode.update once per step (graph_manager.py:2436-2438), float32 state
    (the node's). 256^3 periodic Taylor-Green, and the 128x128x256 sphere
    channel (reference walls, uniform body force (0, 0, 1e-5) as a boundary
    input). Each update is one launch (k_pair<float, local collide, stream +
    moments>); `pair off` is the same loop with a push launch plus a moments
    pass per update. The one-launch kernel moves 152 B of state + 20 B of
    moments per cell (+1 B of mask with walls)."""
    import torch
    from maddening_amd import workloads
    from maddening_amd.graph import GraphManager
    from maddening_amd.node import LBMNode
    out = {}
    n = 256
    rho, u = taylor_green_device(n, 0.05, dev)
    chan = (128, 128, 256)
    mask_np = workloads.sphere_packing(chan, porosity=0.7, rmin=6, rmax=12, seed=0)
    for name, shape, mask, ext in (
            ("taylor_green_256", (n, n, n), None, {}),
            ("sphere_channel_128x128x256", chan, mask_np, {"body_force": [0.0, 0.0, 1e-5]})):
        cells = shape[0] * shape[1] * shape[2]
        rec = {"grid": list(shape), "updates": nsteps}
        for pair in ("auto", "off"):
            node = LBMNode("fluid", 1.0, grid_shape=shape, viscosity=(0.8 - 0.5) / 3.0,
                           wall_mask=mask, device=dev, pair=pair)
            k = node.kernel
            gm = GraphManager()
            gm.add_node(node)
            if mask is None:
                st = gm.get_node_state("fluid")
                st["f"] = k.aos_view(k.equilibrium(rho, u))
                gm.set_node_state("fluid", st)
            gm.compile()
            e = {"fluid": dict(ext)}
            gm.run(3, external_inputs=e)
            torch.cuda.synchronize()
            times = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                gm.run(nsteps, external_inputs=e)
                e1.record()
                torch.cuda.synchronize()
                times.append(e0.elapsed_time(e1) / 1e3 / nsteps)
            t = statistics.median(times)
            r = {"ms_per_update": round(t * 1e3, 4), "MLUPS": round(cells / t / 1e6, 1),
                 "launches_per_update": k.schedule(1)[0]}
            if pair == "auto":
                # the launch the node makes: a mask without walls runs the
                # periodic kernel (LBMNode._kernel_for)
                st = gm.get_node_state("fluid")
                src = node._soa(st["f"])
                kk, m = node._kernel_for(st["wall_mask"])
                kt = time_phase(kk, 6, src, mask=m, moments=True)
                bpc = 152 + 20 + (1 if mask is not None else 0)
                r["kernel_us"] = round(kt * 1e6, 2)
                r["kernel_GBs"] = round(bpc * cells / kt / 1e9, 1)
                r["kernel_bytes_per_cell"] = bpc
                del src, m, st
            rec["one_launch" if pair == "auto" else "push_plus_moments"] = r
            del node, gm, k
            torch.cuda.empty_cache()
        rec["speedup"] = round(rec["push_plus_moments"]["ms_per_update"]
                               / rec["one_launch"]["ms_per_update"], 3)
        out[name] = rec
    return out


def config4_forward_leg(dev, nsteps=500, shape=(128, 128, 256), reps=3):
    """Forward of the sphere-packed channel (porosity 0.7, radii 6-12, seed 0;
    reference walls, tau field 0.7, force (0, 0, 1e-5)), float32: the
    library's choice (general two-step launches: walls and the tau field
    inside the LDS march) against one fused launch per step (pair off); the
    same bits either way."""
    import torch
    from maddening_amd import workloads
    from maddening_amd.lbm import LBMKernel
    mask = torch.from_numpy(workloads.sphere_packing(shape, porosity=0.7, rmin=6, rmax=12,
                                                     seed=0)).to(dev)
    tau = torch.full(shape, 0.7, device=dev)
    res = {}
    outs = {}
    for pair in ("off", "auto"):
        k = LBMKernel(shape, tau=0.7, storage="f32", bb="reference", pair=pair, device=dev)
        f0 = k.equilibrium()
        work = k.workspace(nsteps)
        o = k.empty()
        k.step(f0, 10, mask=mask, force=[0, 0, 1e-5], tau_field=tau, work=work, out=o)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            k.step(f0, nsteps, mask=mask, force=[0, 0, 1e-5], tau_field=tau, work=work, out=o)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / nsteps)
        outs[pair] = o.clone()
        res["single" if pair == "off" else "two_step"] = {
            "ms_per_step": round(statistics.median(times), 4),
            "launches": k.schedule(nsteps, with_fields=True)[0]}
        del k, f0, work, o
        torch.cuda.empty_cache()
    res["same_bits"] = bool(torch.equal(outs["auto"], outs["off"]))
    res["speedup"] = round(res["single"]["ms_per_step"] / res["two_step"]["ms_per_step"], 3)
    return res


def batch_leg(dev, nsteps=200, seed=0):
    """64 lattices (32 D2Q9 edge U[64,1024], 32 D3Q19 edge U[32,160], tau
    U[0.55,1.5], Taylor-Green or lid-driven-cavity init), float32: one
    graph-captured launch sequence (the periodic D3Q19 lattices two steps
    per launch, the library's default) vs one lbm_step call per lattice,
    with the same graph at one step per launch beside it."""
    import torch
    from maddening_amd import workloads
    from maddening_amd.autotune import tune_batch
    from maddening_amd.batch import LatticeBatch
    specs = workloads.batch_specs(seed)
    rng = np.random.default_rng(seed)
    for s in specs:
        if s["init"] == "cavity":
            s["bb"] = "halfway"
            s["mask"] = workloads.cavity_mask(*s["shape"])
            s["lid_velocity"] = (0.1, 0.0, 0.0)
    batch = LatticeBatch(specs, device=dev)
    f0s = []
    for k, s in zip(batch.kernels, specs):
        if s["init"] == "cavity":
            f0s.append(k.equilibrium())
        elif k.D == 2:
            r, u = workloads.taylor_green_2d(*k.grid_shape, float(rng.uniform(0.01, 0.05)))
            f0s.append(k.equilibrium(torch.from_numpy(r).float(), torch.from_numpy(u).float()))
        else:
            r, u = workloads.taylor_green_3d(k.grid_shape[0], float(rng.uniform(0.01, 0.05)))
            f0s.append(k.equilibrium(torch.from_numpy(r).float(), torch.from_numpy(u).float()))
    tune = tune_batch(batch, use_cache=False)
    batch.set_states(f0s)
    h = batch.graph(nsteps, tune["grid"])
    batch.launch_graph(h)  # warm
    torch.cuda.synchronize()
    batch.set_states(f0s)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    batch.launch_graph(h)
    e1.record()
    torch.cuda.synchronize()
    batch.set_states(f0s)
    batch.run_per_lattice(5)  # warm
    torch.cuda.synchronize()
    batch.set_states(f0s)
    e1b = torch.cuda.Event(enable_timing=True)
    e1b.record()
    batch.run_per_lattice(nsteps)
    e2.record()
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) / 1e3
    t_loop = e1b.elapsed_time(e2) / 1e3
    batch.close()
    # the same graph with every lattice one step per launch
    # (lbm_batch_plan_pairs off; bit-identical)
    batch2 = LatticeBatch(specs, device=dev, pairs=False)
    batch2.plan(tune["cells_per_thread"])
    h2 = batch2.graph(nsteps, tune["grid"])
    batch2.set_states(f0s)
    batch2.launch_graph(h2)  # warm
    torch.cuda.synchronize()
    batch2.set_states(f0s)
    e3, e4 = (torch.cuda.Event(enable_timing=True) for _ in range(2))
    e3.record()
    batch2.launch_graph(h2)
    e4.record()
    torch.cuda.synchronize()
    t_pair = e3.elapsed_time(e4) / 1e3
    batch2.close()
    return {"lattices": len(specs), "cells": batch.cells, "steps": nsteps,
            "graph_MLUPS": round(batch.cells * nsteps / t_graph / 1e6, 1),
            "per_lattice_MLUPS": round(batch.cells * nsteps / t_loop / 1e6, 1),
            "speedup": round(t_loop / t_graph, 3),
            "one_step_graph_MLUPS": round(batch.cells * nsteps / t_pair / 1e6, 1),
            "tuned": {"cells_per_thread": tune["cells_per_thread"], "grid": tune["grid"]}}


def d2q9_leg(dev, n=8192, nsteps=200, reps=3):
    """D2Q9 periodic Taylor-Green at n x n (67 M cells at 8192^2): the 2-D
    path's throughput for each storage type. Algorithmic bytes per cell
    update 2 * 9 * s_f (72 B for 32-bit storage, 36 B for 16-bit). 32-bit
    storage runs two steps per launch (k_pair9) by the library's own
    choice; its one-update-per-launch rate is reported beside it."""
    import torch
    from maddening_amd import workloads
    from maddening_amd.lbm import LBMKernel
    rho, u = workloads.taylor_green_2d(n, n, 0.05)
    rho = torch.from_numpy(rho).float().to(dev)
    u = torch.from_numpy(u).float().to(dev)
    out = {"grid": [n, n], "steps": nsteps}
    for storage, pair in (("f32dev", "auto"), ("f32dev", "off"), ("f16dev", "auto")):
        k = LBMKernel((n, n), lattice="D2Q9", tau=0.8, storage=storage, pair=pair, device=dev)
        f0 = k.equilibrium(rho, u)
        work = k.workspace(nsteps)
        res = k.empty()
        k.step(f0, 5, out=res, work=work, moments=False)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            k.step(f0, nsteps, out=res, work=work, moments=False)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / 1e3)
        t = statistics.median(times)
        bpc = 2 * 9 * (2 if storage == "f16dev" else 4)
        rec = {"MLUPS": round(n * n * nsteps / t / 1e6, 1),
               "algorithmic_GBs": round(bpc * n * n * nsteps / t / 1e9, 1),
               "bytes_per_cell_update": bpc, "two_step_launches": k.schedule(nsteps)[1]}
        out[storage if pair == "auto" else storage + "_single_pass"] = rec
        del k, f0, work, res
        torch.cuda.empty_cache()
    return out


def long_rows_leg(dev, shape=(128, 128, 1024), nsteps=200, reps=3):
    """D3Q19 periodic Taylor-Green with rows longer than one workgroup holds
    whole (1024 cells), f32dev: the library's own choice (two steps per
    launch in 64-column tiles) against one step per launch. Same bits either
    way; 152 B per cell update."""
    import numpy as np
    import torch
    from maddening_amd.lbm import LBMKernel
    # Taylor-Green vortex with one period along each axis (k = 2 pi / n per
    # axis), unit density
    X, Y, Z = np.meshgrid(*(2 * np.pi * np.arange(n) / n for n in shape), indexing="ij")
    u = np.stack([0.05 * np.sin(X) * np.cos(Y) * np.cos(Z),
                  -0.05 * np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X)], axis=-1)
    rho = torch.ones(shape, dtype=torch.float32, device=dev)
    u = torch.from_numpy(u).float().to(dev)
    del X, Y, Z
    cells = shape[0] * shape[1] * shape[2]
    out = {"grid": list(shape), "steps": nsteps}
    for pair in ("auto", "off"):
        k = LBMKernel(shape, lattice="D3Q19", tau=0.8, storage="f32dev", pair=pair, device=dev)
        f0 = k.equilibrium(rho, u)
        work = k.workspace(nsteps)
        res = k.empty()
        k.step(f0, 5, out=res, work=work, moments=False)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            k.step(f0, nsteps, out=res, work=work, moments=False)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / 1e3)
        t = statistics.median(times)
        out["MLUPS" if pair == "auto" else "single_pass_MLUPS"] = round(cells * nsteps / t / 1e6, 1)
        if pair == "auto":
            out["two_step_launches"] = k.schedule(nsteps)[1]
            out["algorithmic_GBs"] = round(152 * cells * nsteps / t / 1e9, 1)
        del k, f0, work, res
        torch.cuda.empty_cache()
    out["speedup"] = round(out["MLUPS"] / out["single_pass_MLUPS"], 3)
    return out


def pipe_leg(dev, shape=(256, 128, 128), nsteps=200, reps=3):
    """LBMPipeNode (lbm_pipe.py:751-879) on the GPU: single phase with the
    D3Q7 tracer, and Shan-Chen multiphase, float32, propeller + gravity.
    Algorithmic bytes per cell and step: single phase 152 (D3Q19) + 56
    (D3Q7) + 1 (flags) = 209; multiphase, two pull passes: 76 gather + 4 psi
    + 4 floor scale + 1 flags, then 76 gather + 4 scale + 4 psi + 76 write
    + 1 flags = 246."""
    import torch
    from maddening_amd.pipe import LBMPipeNode
    nx, ny, nz = shape
    cells = nx * ny * nz
    out = {"grid": list(shape), "steps": nsteps}
    for name, kw, bpc in (("single_phase", dict(fill_fraction=0.75, gravity=-1e-4), 209),
                          ("multiphase", dict(fill_fraction=0.55, gravity=-1e-3, G=-4.5), 246)):
        node = LBMPipeNode("fluid", 0.01, nx=nx, ny=ny, nz=nz, propeller_x=nx // 6,
                           propeller_strength=0.001, device=dev, **kw)
        st = node.initial_state()
        node.run(st, {}, 5)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            node.run(st, {}, nsteps)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) / 1e3)
        t = statistics.median(times)
        out[name] = {"MLUPS": round(cells * nsteps / t / 1e6, 1),
                     "GB/s": round(bpc * cells * nsteps / t / 1e9, 1),
                     "ms_per_step": round(t / nsteps * 1e3, 4), "bytes_per_cell": bpc}
        del node, st
        torch.cuda.empty_cache()
    return out


def read_traffic(storage, pair=False):
    """HBM bytes per launch of the fused kernel (pair: of the two-step kernel)
    from the committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled
    by the copy-kernel factor), and where that figure comes from. Counters
    need a run under rocprofv3 --pmc, so they cannot be read inside this
    timed run."""
    name = f"pmc_d3q19_256_{storage}_pair.json" if pair else f"pmc_d3q19_256_{storage}.json"
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None, None
    src = f"profiles/{name}"
    if rec.get("date"):
        src += f" ({rec['date']})"
    return rec.get("traffic_bytes_per_launch"), src


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
    ap.add_argument("--storage", default="f32dev",
                    help="population storage of the headline run: f32dev (float32 f - w, "
                         "the default: same bytes as f32, 10x smaller rounding) or f32")
    ap.add_argument("--variants", default="f32,f16dev,bf16dev",
                    help="extra storage types to time (comma list, '' for none)")
    ap.add_argument("--legs",
                    default="precision,adjoint,node_update,config4_forward,batch,pipe,d2q9,long_rows",
                    help="extra measurements: precision (storage types vs float64 at "
                         "100/1000/5000 steps), adjoint (forward+adjoint ms/step, gradient "
                         "vs central differences), batch (64 uneven lattices), pipe "
                         "(LBMPipeNode single phase and multiphase), d2q9 (8192^2 D2Q9), "
                         "long_rows (D3Q19 128x128x1024, rows in 64-column tiles), "
                         "node_update (GraphManager.step with LBMNode.update, one launch per "
                         "update), config4_forward (the sphere channel one or two steps per "
                         "launch)")
    ap.add_argument("--pair", default="auto", choices=["auto", "off", "force"],
                    help="two pull steps per launch (LBM_FLAG_PAIR_*)")
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
                         xcd_remap=args.xcd_remap, pair=args.pair, device=dev)

    if args.calibrate_copy:
        a = torch.ones(1 << 28, device=dev)  # 1 GiB float32
        b = torch.empty_like(a)
        for _ in range(5):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b

    kern = make(args.storage)
    f0 = kern.equilibrium(rho, u)
    # Dominant kernel first: the two-step launch, timed on its own (two cell
    # updates per cell, the state read and written once). After an idle
    # second the device needs some 20 ms of load to reach the clocks a long
    # run holds (a 256^3 20-step call measured 0.306 ms/step right after a
    # 5-step warm-up, 0.276 after 60 ms of launches, 0.314 after 1 s idle),
    # so these 100 launches also bring the K-step call below to those clocks.
    _, pairs0 = kern.schedule(args.steps)
    per_launch_s = time_pair_launch(kern, f0, reps=100) if pairs0 else None
    ev, wall, launches, pairs, out = time_kernel(kern, f0, args.steps, args.warmup)
    mlups = cells * args.steps / ev / 1e6
    b_cell = 2 * 19 * STORAGE_BYTES[args.storage]
    rv = auto_row_vec(args.storage, n, args.row_vec) if kern.scheme != "push" else 0
    if not pairs:
        per_launch_s = ev / launches
    updates = 2 if pairs else 1
    # Algorithmic HBM bytes of one launch: the state read once and written
    # once, b_cell per cell, whatever the number of updates the launch makes
    # (a two-step launch keeps its intermediate state in LDS). The update
    # rate the launch sustains, priced at b_cell per update, is reported
    # beside it as update_frac.
    achieved = b_cell * cells / per_launch_s / 1e9
    update_rate = updates * b_cell * cells / per_launch_s / 1e9
    traffic, traffic_src = read_traffic(args.storage, pair=bool(pairs))
    storage_dtype = {"f32": "f32", "f32dev": "f32dev", "f64": "f64", "f16dev": "f16dev",
                     "bf16dev": "bf16dev"}[args.storage]
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
        "dtype": storage_dtype,
        "data": "synthetic",
        "config": {"workload": f"D3Q19 BGK periodic {n}^3 Taylor-Green vortex, u0=0.05, "
                               f"tau=0.8, populations stored as {args.storage}",
                   "cells": cells, "scheme": kern.scheme, "launches": launches,
                   "two_step_launches": pairs,
                   # the first launch runs three units (an even K, LBM_FLAG_TRI_OFF)
                   "three_unit_launches": int(bool(pairs) and args.steps % 2 == 0 and
                                              args.steps >= 4 and launches == args.steps // 2),
                   "compute_dtype": "f64" if args.storage == "f64"
                   else "f32", "wall_s_timed_call": round(wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": f"{fused_kernel_label(args.storage, rv, pairs)} "
                               f"({b_cell} B algorithmic per cell and launch, "
                               f"{updates} update(s) per cell per launch)",
                     "avg_launch_us": round(per_launch_s * 1e6, 2),
                     "updates_per_launch": updates,
                     # HBM bytes the launch actually moves (rocprofv3
                     # FETCH_SIZE + WRITE_SIZE) over its live duration
                     "hbm_GBs": (round(traffic / per_launch_s / 1e9, 1)
                                 if traffic else None),
                     "hbm_frac": (round(traffic / per_launch_s / 1e9 / HBM_PEAK_GBS, 4)
                                  if traffic else None),
                     # the update rate priced at b_cell per cell update
                     "update_GBs": round(update_rate, 1),
                     "update_frac": round(update_rate / HBM_PEAK_GBS, 4)},
    }
    if pairs:
        # the same workload with one update per launch (LBM_FLAG_PAIR_OFF)
        ks = LBMKernel((n, n, n), lattice="D3Q19", tau=0.8, storage=args.storage,
                       scheme=args.scheme, block_x=args.block_x, row_vec=args.row_vec,
                       xcd_remap=args.xcd_remap, pair="off", device=dev)
        ev1, _, launches1, _, out1 = time_kernel(ks, f0, args.steps, args.warmup)
        t1 = ev1 / launches1
        result["single_pass"] = {
            "MLUPS": round(cells * args.steps / ev1 / 1e6, 1),
            "kernel": f"{fused_kernel_label(args.storage, rv)}",
            "avg_launch_us": round(t1 * 1e6, 2),
            "achieved_GBs": round(b_cell * cells / t1 / 1e9, 1),
            "frac": round(b_cell * cells / t1 / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(args.storage)[0],
            "same_bits_as_two_step": bool(torch.equal(out1, out))}
        del out1, ks
    extra = {}
    for storage in [s for s in args.variants.split(",") if s]:
        k2 = make(storage)
        f2 = k2.equilibrium(rho, u)
        # same order as the headline: the variant's own two-step launch
        # first, which also brings the device back to sustained clocks
        _, p2 = k2.schedule(args.steps)
        t2 = time_pair_launch(k2, f2, reps=100) if p2 else None
        ev2, _, launches2, pairs2, _ = time_kernel(k2, f2, args.steps, args.warmup)
        bc2 = 2 * 19 * STORAGE_BYTES[storage]
        extra[storage] = {"MLUPS": round(cells * args.steps / ev2 / 1e6, 1),
                          "algorithmic_GBs": round(bc2 * cells * args.steps / ev2 / 1e9, 1),
                          "two_step_launches": pairs2,
                          "speedup_vs_headline": round(ev / ev2, 3), "_ev": ev2}
        if t2:
            extra[storage]["pair_launch_us"] = round(t2 * 1e6, 2)
            extra[storage]["pair_launch_GBs"] = round(bc2 * cells / t2 / 1e9, 1)
        if pairs2:
            ks = LBMKernel((n, n, n), lattice="D3Q19", tau=0.8, storage=storage,
                           scheme=args.scheme, pair="off", device=dev)
            ev3, _, _, _, _ = time_kernel(ks, f2, args.steps, args.warmup)
            extra[storage]["single_pass_MLUPS"] = round(cells * args.steps / ev3 / 1e6, 1)
            extra[storage]["_ev1"] = ev3
            del ks
        del f2, k2
    ref32 = extra["f32"]["_ev"] if "f32" in extra else (ev if args.storage == "f32" else None)
    ref32_1 = extra["f32"].get("_ev1") if "f32" in extra else None
    for v in extra.values():
        ev_v = v.pop("_ev")
        v.pop("_ev1", None)
        if ref32 is not None:
            v["speedup_vs_f32"] = round(ref32 / ev_v, 3)
        if ref32_1 is not None:
            # against float32 storage with one update per launch
            v["speedup_vs_f32_single_pass"] = round(ref32_1 / ev_v, 3)
    if extra:
        result["variants"] = extra
    del out, f0
    torch.cuda.empty_cache()
    legs = [s for s in args.legs.split(",") if s]
    if "precision" in legs:
        result["precision"] = precision_leg(n, dev)
    if "adjoint" in legs:
        result["adjoint"] = adjoint_leg(dev)
    if "node_update" in legs:
        result["node_update"] = node_update_leg(dev)
    if "config4_forward" in legs:
        result["config4_forward"] = config4_forward_leg(dev)
    if "batch" in legs:
        result["batch"] = batch_leg(dev)
    if "pipe" in legs:
        result["pipe"] = pipe_leg(dev)
    if "d2q9" in legs:
        result["d2q9"] = d2q9_leg(dev)
    if "long_rows" in legs:
        result["long_rows"] = long_rows_leg(dev)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
