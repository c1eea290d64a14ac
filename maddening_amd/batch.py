"""Batches of independent lattices advanced by one launch per step.

``LatticeBatch`` owns two float32 SoA state buffers per lattice and the
device plan of ``lbm_batch_plan``. ``run(nsteps)`` launches the whole batch
(nsteps + 1 launches); ``graph(nsteps)`` captures those launches once into a
HIP graph that ``launch_graph`` replays. ``run_per_lattice`` is the
comparison point: one ``lbm_step`` call per lattice.

The batch buffers pad each population plane to a multiple of
``plane_align`` cells (64 by default), so that every store of a step is an
aligned 256-byte run whatever the lattice's cell count. Measured on MI355X,
a lattice whose cell count is not a multiple of 64 otherwise stores
misaligned runs and its step moves about a third fewer bytes per second.
``state_views`` gives the padded buffers as (Q, *G) views; ``states`` and
``set_states`` use the ordinary lbm_step layout. Lattices under halfway
bounce-back also get a buffer of per-cell link words,
rebuilt from the mask at the start of every run, so that their steps read
one word per cell instead of the neighbours' mask bytes (8 bytes per cell:
the link word, and the moving-wall links read only next to a moving wall).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .lbm import LBMKernel, _ptr, _stream


class LatticeBatch:
    """Lattices of any size and type stepped together.

    ``specs``: sequence of dicts with keys ``lattice`` ("D3Q19"/"D2Q9"),
    ``shape``, ``tau``, optional ``force`` (D,), ``bb`` ("none",
    "reference", "halfway"), ``mask`` (uint8 array), ``lid_velocity``.
    ``pairs``: periodic D3Q19 lattices with rows of at most 341 cells run
    two steps per launch (same bits; ``lbm_batch_plan_pairs``). On by
    default: the 64-lattice benchmark batch ran at 44.8k MLUPS with it
    against 40.5k without.
    """

    def __init__(self, specs: Sequence[dict], *, cells_per_thread: int = 1, device="cuda",
                 plane_align: int = 64, link_words: bool = True, pairs: bool = True):
        self.device = torch.device(device)
        self.pairs = bool(pairs)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        if plane_align < 1:
            raise ValueError("plane_align must be >= 1")
        self.kernels = []
        self.buffers = []      # (Q, *G) views of the padded state buffers
        self.pitches = []      # cells from one population plane to the next
        self.masks = []
        self.specs = list(specs)
        self._std = None       # lbm_step-layout buffers of the per-lattice baseline
        self.links = []        # link-word buffers of halfway lattices (None elsewhere)
        for s in self.specs:
            bb = s.get("bb", "none")
            k = LBMKernel(s["shape"], lattice=s["lattice"], tau=s["tau"], storage="f32", bb=bb,
                          lid_velocity=s.get("lid_velocity", (0.0, 0.0, 0.0)), device=self.device)
            self.kernels.append(k)
            pitch = -(-k.ncells // plane_align) * plane_align
            self.pitches.append(pitch)
            self.buffers.append(tuple(
                torch.empty(k.Q * pitch, dtype=torch.float32, device=self.device)
                .view(k.Q, pitch)[:, :k.ncells].view(k.state_shape) for _ in range(2)))
            self.links.append(torch.empty(2 * k.ncells, dtype=torch.int32, device=self.device)
                              if (bb == "halfway" and link_words) else None)
            m = s.get("mask")
            self.masks.append(None if m is None else
                              torch.as_tensor(np.asarray(m, np.uint8), device=self.device))
        self.cells = sum(k.ncells for k in self.kernels)
        self._plan = None
        self._handle = None
        self._graphs = []
        self.cells_per_thread = 0
        self.total_tiles = 0
        self.plan(cells_per_thread)

    def _items(self):
        arr = (_lib.LbmBatchItem * len(self.kernels))()
        for i, (k, s) in enumerate(zip(self.kernels, self.specs)):
            it = arr[i]
            it.q = k.Q
            it.bb_mode = _lib.BB_NAMES[k.bb]
            dims = list(k.grid_shape) + [1] * (3 - k.D)
            for j in range(3):
                it.dims[j] = dims[j]
            it.tau = k.tau
            for j, v in enumerate(s.get("force", (0.0,) * k.D)):
                it.force[j] = float(v)
            for j in range(3):
                it.lid_velocity[j] = k.lid_velocity[j]
            it.f[0] = self.buffers[i][0].data_ptr()
            it.f[1] = self.buffers[i][1].data_ptr()
            it.mask = None if self.masks[i] is None else self.masks[i].data_ptr()
            it.pitch = self.pitches[i]
            it.links = None if self.links[i] is None else self.links[i].data_ptr()
        return arr

    def plan(self, cells_per_thread: int) -> None:
        """(Re)build the device plan with tiles of 256 * cells_per_thread cells.
        Graphs captured from the previous plan are released."""
        self.close()
        L = _lib.lib()
        n = len(self.kernels)
        plan = torch.empty(L.lbm_batch_plan_size(n), dtype=torch.uint8, device=self.device)
        tiles = ctypes.c_longlong(0)
        handle = ctypes.c_void_p()
        items = self._items()
        _lib.check(L.lbm_batch_plan_create(_stream(self.device), items, n, int(cells_per_thread),
                                           _ptr(plan), ctypes.byref(handle), ctypes.byref(tiles)),
                   "lbm_batch_plan_create")
        _lib.check(L.lbm_batch_plan_pairs(handle, int(self.pairs)), "lbm_batch_plan_pairs")
        self._plan = plan
        self._handle = handle
        self.cells_per_thread = int(cells_per_thread)
        self.total_tiles = int(tiles.value)
        # a grid tuned for the previous tiling does not carry over
        self.grid = None

    def default_grid(self) -> int:
        """The grid `run` and `graph` use when none is given: the one
        autotune.tune_batch picked for this tiling, else 8 workgroups per CU."""
        if getattr(self, "grid", None):
            return int(min(self.total_tiles, self.grid))
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        return int(min(self.total_tiles, 8 * cus))

    # -- state ---------------------------------------------------------------
    def set_states(self, states: Sequence[torch.Tensor]) -> None:
        """Copy initial SoA f32 states (lbm_step layout) into the first buffer
        of each lattice, and into the per-lattice baseline's buffers."""
        for i, ((a, _), f) in enumerate(zip(self.buffers, states)):
            a.copy_(f)
            if self._std is not None:
                self._std[i][0].copy_(f)

    def state_views(self, nsteps: int):
        """(Q, *G) views of the padded buffers holding the final state after a
        run of ``nsteps`` steps (not contiguous when planes are padded)."""
        return [bufs[(nsteps + 1) % 2] for bufs in self.buffers]

    def states(self, nsteps: int):
        """The final states after a run of ``nsteps`` steps, in the lbm_step
        layout (copies where planes are padded)."""
        return [v if v.is_contiguous() else v.contiguous() for v in self.state_views(nsteps)]

    # -- execution -------------------------------------------------------------
    def run(self, nsteps: int, grid: Optional[int] = None) -> None:
        g = grid or self.default_grid()
        _lib.check(_lib.lib().lbm_batch_run(_stream(self.device), self._handle, int(nsteps),
                                            int(g)), "lbm_batch_run")

    def graph(self, nsteps: int, grid: Optional[int] = None) -> ctypes.c_void_p:
        """Capture a run of ``nsteps`` into a HIP graph (freed with the batch)."""
        g = grid or self.default_grid()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lbm_batch_graph_create(self._handle, int(nsteps), int(g),
                                                     ctypes.byref(h)), "lbm_batch_graph_create")
        self._graphs.append(h)
        return h

    def launch_graph(self, h: ctypes.c_void_p) -> None:
        _lib.check(_lib.lib().lbm_batch_graph_launch(_stream(self.device), h),
                   "lbm_batch_graph_launch")

    def run_per_lattice(self, nsteps: int):
        """Baseline: one lbm_step call per lattice (nsteps + 1 launches each),
        from the states last given to set_states, in lbm_step-layout buffers
        of its own (made from the batch's first buffers on the first call).
        Returns the final states."""
        if self._std is None:
            self._std = [(a.clone(memory_format=torch.contiguous_format),
                          torch.empty_like(a, memory_format=torch.contiguous_format))
                         for a, _ in self.buffers]
        outs = []
        for k, (a, b), m, s in zip(self.kernels, self._std, self.masks, self.specs):
            k.step(a, nsteps, mask=m, force=list(s.get("force", (0.0,) * k.D)), out=b,
                   moments=False)
            outs.append(b)
        return outs

    def close(self) -> None:
        """Release captured graphs and the plan handle."""
        for h in self._graphs:
            _lib.lib().lbm_batch_graph_destroy(h)
        self._graphs = []
        if self._handle is not None:
            _lib.lib().lbm_batch_plan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
