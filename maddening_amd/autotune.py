"""Launch-shape autotuner with an on-disk cache.

For a single lattice it times the one-cell-per-thread kernels at each
workgroup width and the row-per-wave kernels at each legal vector width;
for a batch it times tile sizes (cells per thread) against persistent grid
sizes. Each candidate runs a few warm steps, then a timed short run; the
fastest is stored in a JSON cache keyed by device name and problem shape
(``$MADDENING_AMD_TUNE_CACHE``, default ``~/.cache/maddening_amd/tune.json``).
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Optional

import torch

from .lbm import LBMKernel

_CACHE = os.environ.get("MADDENING_AMD_TUNE_CACHE",
                        os.path.join(os.path.expanduser("~"), ".cache", "maddening_amd",
                                     "tune.json"))


def _load() -> dict:
    try:
        with open(_CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    data = _load()
    data[key] = value
    try:
        os.makedirs(os.path.dirname(_CACHE), exist_ok=True)
        with open(_CACHE, "w") as fh:
            json.dump(data, fh, indent=1)
    except OSError:
        pass  # a read-only home only loses the cache


def _time(fn, reps: int) -> float:
    """Best of ``reps`` timed calls (after one untimed call), in ms: the
    minimum is the estimate least disturbed by clock and queue noise."""
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def tune_kernel(kern: LBMKernel, nsteps: int = 8, reps: int = 3, use_cache: bool = True) -> dict:
    """Pick block_x / row_vec / scheme for ``kern``'s lattice; applies and
    returns the winner with every candidate's ms per step."""
    dev = torch.cuda.get_device_name(kern.device)
    key = f"kernel|{dev}|{kern.lat.name}|{kern.grid_shape}|{kern.storage}|{kern.bb}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            kern.block_x, kern.row_vec, kern.scheme = hit["block_x"], hit["row_vec"], hit["scheme"]
            return hit
    cands = [(bx, -1, sc) for bx in (64, 128, 256) for sc in ("pull", "push")]
    if kern.bb == "none":
        cands += [(0, v, "pull") for v in (1, 2, 4)
                  if kern.grid_shape[-1] % (64 * v) == 0 and not (v == 4 and kern.storage == "f64")]
    f = kern.equilibrium()
    out = kern.empty()
    mask = torch.zeros(kern.grid_shape, dtype=torch.uint8, device=kern.device) \
        if kern.bb != "none" else None
    saved = (kern.block_x, kern.row_vec, kern.scheme)
    results = []
    for bx, rv, sc in cands:
        kern.block_x, kern.row_vec, kern.scheme = bx, rv, sc
        work = kern.workspace(nsteps)
        ms = _time(lambda: kern.step(f, nsteps, mask=mask, out=out, work=work, moments=False),
                   reps) / nsteps
        results.append({"block_x": bx, "row_vec": rv, "scheme": sc, "ms_per_step": ms})
    kern.block_x, kern.row_vec, kern.scheme = saved
    best = min(results, key=lambda r: r["ms_per_step"])
    kern.block_x, kern.row_vec, kern.scheme = best["block_x"], best["row_vec"], best["scheme"]
    rec = dict(best, candidates=results)
    _store(key, rec)
    return rec


def tune_batch(batch, nsteps: int = 10, reps: int = 3, use_cache: bool = True,
               grids: Optional[list] = None) -> dict:
    """Pick cells_per_thread (tile = 256 * cpt cells) and the persistent grid
    size for a LatticeBatch; re-plans the batch with the winner."""
    dev = torch.cuda.get_device_name(batch.device)
    sig = [(k.lat.name, k.grid_shape, k.bb) for k in batch.kernels]
    # a digest that is the same in every process (str hashes are salted)
    digest = hashlib.sha1(str(sig).encode()).hexdigest()[:12]
    key = f"batch|{dev}|{digest}|{batch.cells}"
    if use_cache:
        hit = _load().get(key)
        if hit:
            batch.plan(hit["cells_per_thread"])
            return hit
    cus = torch.cuda.get_device_properties(batch.device).multi_processor_count
    results = []
    for cpt in (1, 2, 4):
        batch.plan(cpt)
        gs = grids or sorted({min(batch.total_tiles, m * cus) for m in (2, 4, 8, 16, 32)}
                             | {batch.total_tiles})
        for g in gs:
            ms = _time(lambda: batch.run(nsteps, g), reps) / (nsteps + 1)
            results.append({"cells_per_thread": cpt, "grid": int(g), "ms_per_launch": ms})
    best = min(results, key=lambda r: r["ms_per_launch"])
    batch.plan(best["cells_per_thread"])
    rec = dict(best, candidates=results)
    _store(key, rec)
    return rec
