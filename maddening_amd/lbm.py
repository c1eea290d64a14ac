"""PyTorch-ROCm front-end of the HIP LBM kernels.

Tensors carry device memory and streams; every computation is a call into
libmaddening_lbm.so through its C ABI (include/lbm_c_api.h).

States are SoA tensors of shape (Q, *grid_shape) in the storage dtype:
"f32" / "f64" hold f itself; "f32dev", "f16dev" and "bf16dev" hold f - w
(float32, float16, bfloat16). ``aos_view``
gives the reference's (*grid_shape, Q) view of an f32/f64 state without a
copy.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .lattice import LatticeDescriptor, d2q9, d3q19

_TORCH_DTYPE = {
    _lib.LBM_F32: torch.float32,
    _lib.LBM_F64: torch.float64,
    _lib.LBM_F16_DEV: torch.float16,
    _lib.LBM_BF16_DEV: torch.bfloat16,
    _lib.LBM_F32_DEV: torch.float32,
}


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device: torch.device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


_PAIR_FLAGS = {"auto": 0, "off": _lib.LBM_FLAG_PAIR_OFF, "force": _lib.LBM_FLAG_PAIR_FORCE}


class LBMKernel:
    """A configured lattice: grid, velocity set, relaxation time, storage
    type, wall rule and launch shape.

    Parameters
    ----------
    grid_shape : (n0, n1, n2) for D3Q19 or (n0, n1) for D2Q9.
    lattice : "D3Q19" or "D2Q9".
    tau : BGK relaxation time (> 0.5). Alternatively ``viscosity`` with
        tau = 0.5 + viscosity / cs2 (lbm.py:634).
    storage : "f32", "f64", "f32dev", "f16dev" or "bf16dev".
    bb : "none" (periodic), "reference" (lbm.py:944-948) or "halfway".
    lid_velocity : wall velocity of mask==2 cells under "halfway".
    scheme : "auto", "pull" or "push".
    block_x, row_vec, xcd_remap, xcd_group : launch shape (see
        include/lbm_c_api.h); 0 / 0 / False / 0 pick the measured defaults.
    pair : "auto", "off" or "force": two pull steps per launch
        (LBM_FLAG_PAIR_*); "auto" uses them where they measured faster.
        Results are bit-identical in every mode.
    tri : False sets LBM_FLAG_TRI_OFF (no three-unit first launch in even
        sequences of 32- and 16-bit storage); the same bits either way.
    """

    def __init__(self, grid_shape: Sequence[int], lattice: str = "D3Q19",
                 tau: Optional[float] = None, viscosity: Optional[float] = None,
                 storage: str = "f32", bb: str = "none",
                 lid_velocity: Sequence[float] = (0.0, 0.0, 0.0),
                 scheme: str = "auto", block_x: int = 0, row_vec: int = 0,
                 xcd_remap: bool = False, xcd_group: int = 0, pair: str = "auto",
                 tri: bool = True, device="cuda"):
        if lattice.upper() == "D3Q19":
            self.lat: LatticeDescriptor = d3q19()
        elif lattice.upper() == "D2Q9":
            self.lat = d2q9()
        else:
            raise ValueError(f"Unknown lattice '{lattice}'. Supported: 'D3Q19', 'D2Q9'.")
        grid_shape = tuple(int(n) for n in grid_shape)
        if len(grid_shape) != self.lat.D:
            raise ValueError(f"grid_shape has {len(grid_shape)} dims but lattice "
                             f"{self.lat.name} requires {self.lat.D} dims.")
        if tau is None:
            if viscosity is None:
                raise ValueError("give tau or viscosity")
            tau = 0.5 + viscosity / self.lat.cs2
        if not tau > 0.5:
            raise ValueError(f"tau must be > 0.5 (got {tau}). Increase viscosity.")
        if storage not in _lib.STORAGE_NAMES:
            raise ValueError(f"storage must be one of {list(_lib.STORAGE_NAMES)}")
        if bb not in _lib.BB_NAMES:
            raise ValueError(f"bb must be one of {list(_lib.BB_NAMES)}")
        if scheme not in _lib.SCHEME_NAMES:
            raise ValueError(f"scheme must be one of {list(_lib.SCHEME_NAMES)}")
        if pair not in _PAIR_FLAGS:
            raise ValueError(f"pair must be one of {list(_PAIR_FLAGS)}")
        self.pair = pair
        self.grid_shape = grid_shape
        self.D, self.Q = self.lat.D, self.lat.Q
        self.tau = float(tau)
        self.storage = storage
        self.bb = bb
        self.lid_velocity = tuple(float(v) for v in lid_velocity) + (0.0,) * 3
        self.scheme = scheme
        self.block_x = int(block_x)
        self.xcd_remap = bool(xcd_remap)
        self.xcd_group = int(xcd_group)
        self.tri = bool(tri)
        self.row_vec = int(row_vec)
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype = _TORCH_DTYPE[_lib.STORAGE_NAMES[storage]]
        self.ncells = int(np.prod(grid_shape))
        _lib.lib()  # fail loudly now if the library is missing

    # -- descriptors -----------------------------------------------------
    FACES = ("x_min", "x_max", "y_min", "y_max", "z_min", "z_max")

    def desc(self, force=None, tau: Optional[float] = None, faces=None) -> _lib.LbmDesc:
        """Descriptor for a call. ``faces``: optional ((inlet_face, inlet_rho),
        (outlet_face, outlet_rho)) with face names as in ``FACES`` or None."""
        d = _lib.LbmDesc()
        d.q = self.Q
        d.storage = _lib.STORAGE_NAMES[self.storage]
        d.bb_mode = _lib.BB_NAMES[self.bb]
        d.scheme = _lib.SCHEME_NAMES[self.scheme]
        dims = list(self.grid_shape) + [1] * (3 - self.D)
        for i in range(3):
            d.dims[i] = dims[i]
        d.tau = self.tau if tau is None else float(tau)
        if force is not None and not isinstance(force, torch.Tensor):
            fv = [float(v) for v in force]
            if len(fv) != self.D:
                raise ValueError(f"uniform body_force must have shape ({self.D},)")
            for i, v in enumerate(fv):
                d.force[i] = v
        for i in range(3):
            d.lid_velocity[i] = self.lid_velocity[i]
        d.block_x = self.block_x
        d.row_vec = self.row_vec
        d.flags = ((_lib.LBM_FLAG_XCD_REMAP if self.xcd_remap else 0) | _PAIR_FLAGS[self.pair]
                   | (0 if self.tri else _lib.LBM_FLAG_TRI_OFF))
        d.xcd_group = self.xcd_group
        if faces is not None:
            for slot, spec in zip(("inlet", "outlet"), faces):
                if spec is None:
                    continue
                name, rho = spec
                if name not in self.FACES or self.FACES.index(name) // 2 >= self.D:
                    raise ValueError(f"unknown or out-of-lattice face '{name}'")
                setattr(d, f"{slot}_face", self.FACES.index(name) + 1)
                setattr(d, f"{slot}_rho", float(rho))
        return d

    @property
    def state_shape(self):
        return (self.Q,) + self.grid_shape

    def empty(self) -> torch.Tensor:
        return torch.empty(self.state_shape, dtype=self.dtype, device=self.device)

    def schedule(self, nsteps: int, with_fields: bool = False, force=None) -> tuple[int, int]:
        """(launches, launches of two or three units) of ``step(f, nsteps)``;
        with_fields: the call passes a force or tau field; force: the call's
        uniform force (a force rules out the three-unit launch)."""
        nl, npair = ctypes.c_int(0), ctypes.c_int(0)
        if isinstance(force, torch.Tensor):
            force = None if force.dim() > 1 else force.tolist()
        d = self.desc(force=force)
        _lib.check(_lib.lib().lbm_step_schedule(ctypes.byref(d), int(nsteps),
                                                int(bool(with_fields)), ctypes.byref(nl),
                                                ctypes.byref(npair)), "lbm_step_schedule")
        return nl.value, npair.value

    def workspace(self, nsteps: int) -> Optional[torch.Tensor]:
        n = _lib.lib().lbm_workspace_size(ctypes.byref(self.desc()), int(nsteps))
        if n == 0:
            return None
        return torch.empty(n, dtype=torch.uint8, device=self.device)

    # -- argument plumbing ----------------------------------------------
    def _field(self, force):
        if isinstance(force, torch.Tensor):
            if force.shape == (self.D,):
                return None, [float(v) for v in force.tolist()]
            if tuple(force.shape) != self.grid_shape + (self.D,):
                raise ValueError(f"body_force field must have shape {self.grid_shape + (self.D,)}")
            return force.to(device=self.device, dtype=torch.float32).contiguous(), None
        return None, force

    def _mask(self, mask):
        if mask is None:
            return None
        if tuple(mask.shape) != self.grid_shape:
            raise ValueError(f"wall_mask shape {tuple(mask.shape)} != grid_shape {self.grid_shape}")
        return mask.to(device=self.device, dtype=torch.uint8).contiguous()

    def _check_state(self, f):
        if tuple(f.shape) != self.state_shape or f.dtype != self.dtype or not f.is_contiguous():
            raise ValueError(f"state must be a contiguous {self.dtype} tensor of shape "
                             f"{self.state_shape}, got {f.dtype} {tuple(f.shape)}")
        if f.device != self.device:
            raise ValueError(f"state is on {f.device}, kernel on {self.device}")

    # -- compute ----------------------------------------------------------
    def equilibrium(self, rho: Optional[torch.Tensor] = None,
                    u: Optional[torch.Tensor] = None) -> torch.Tensor:
        """f = feq(rho, u) (lbm.py:180-206) as a state."""
        out = self.empty()
        r = None if rho is None else rho.to(self.device, torch.float32).contiguous()
        v = None if u is None else u.to(self.device, torch.float32).contiguous()
        if r is not None and tuple(r.shape) != self.grid_shape:
            raise ValueError("rho must have grid_shape")
        if v is not None and tuple(v.shape) != self.grid_shape + (self.D,):
            raise ValueError("u must have shape grid_shape + (D,)")
        _lib.check(_lib.lib().lbm_equilibrium(_stream(self.device), ctypes.byref(self.desc()),
                                              _ptr(r), _ptr(v), _ptr(out)), "lbm_equilibrium")
        return out

    def step(self, f: torch.Tensor, nsteps: int = 1, *, mask=None, force=None,
             tau_field: Optional[torch.Tensor] = None, moments: bool = True,
             out: Optional[torch.Tensor] = None, work: Optional[torch.Tensor] = None,
             faces=None):
        """Advance ``nsteps`` LBMNode.update steps (lbm.py:899-986).

        ``faces``: optional Zou-He pressure faces ((inlet, rho_in), (outlet,
        rho_out)), either entry None (lbm.py:950-968). Returns (f_new, rho,
        u, p); the moments are None when moments=False.
        """
        self._check_state(f)
        ff, fu = self._field(force)
        m = self._mask(mask)
        if self.bb != "none" and m is None:
            raise ValueError(f"bb='{self.bb}' needs a wall mask")
        tf = None
        if tau_field is not None:
            if tuple(tau_field.shape) != self.grid_shape:
                raise ValueError("tau_field must have grid_shape")
            tf = tau_field.to(self.device, torch.float32).contiguous()
        if out is None:
            out = self.empty()
        else:
            self._check_state(out)
        d = self.desc(force=fu, faces=faces)
        if work is None:
            work = self.workspace(nsteps)
        rho = u = p = None
        if moments:
            rho = torch.empty(self.grid_shape, dtype=torch.float32, device=self.device)
            u = torch.empty(self.grid_shape + (self.D,), dtype=torch.float32, device=self.device)
            p = torch.empty(self.grid_shape, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().lbm_step(_stream(self.device), ctypes.byref(d), _ptr(f), _ptr(out),
                                       _ptr(work), _ptr(m), _ptr(ff), _ptr(tf), _ptr(rho),
                                       _ptr(u), _ptr(p), int(nsteps)), "lbm_step")
        return out, rho, u, p

    def moments(self, f: torch.Tensor, *, mask=None, force=None):
        """(rho, u, p) of a state (lbm.py:209-239, 970-976)."""
        self._check_state(f)
        ff, fu = self._field(force)
        m = self._mask(mask)
        rho = torch.empty(self.grid_shape, dtype=torch.float32, device=self.device)
        u = torch.empty(self.grid_shape + (self.D,), dtype=torch.float32, device=self.device)
        p = torch.empty(self.grid_shape, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().lbm_moments(_stream(self.device), ctypes.byref(self.desc(force=fu)),
                                          _ptr(f), _ptr(m), _ptr(ff), _ptr(rho), _ptr(u), _ptr(p)),
                   "lbm_moments")
        return rho, u, p

    def from_aos(self, f_aos: torch.Tensor) -> torch.Tensor:
        """(*G, Q) float32/float64 array (the reference's layout) -> state."""
        if tuple(f_aos.shape) != self.grid_shape + (self.Q,):
            raise ValueError(f"expected shape {self.grid_shape + (self.Q,)}")
        src = f_aos.to(self.device)
        f64 = src.dtype == torch.float64
        if not f64:
            src = src.to(torch.float32)
        src = src.contiguous()
        out = self.empty()
        _lib.check(_lib.lib().lbm_convert(_stream(self.device), ctypes.byref(self.desc()), 1,
                                          int(f64), _ptr(src), _ptr(out)), "lbm_convert")
        return out

    def to_aos(self, f: torch.Tensor, dtype=torch.float32) -> torch.Tensor:
        """state -> (*G, Q) float32/float64 array (the reference's layout)."""
        self._check_state(f)
        out = torch.empty(self.grid_shape + (self.Q,), dtype=dtype, device=self.device)
        _lib.check(_lib.lib().lbm_convert(_stream(self.device), ctypes.byref(self.desc()), 0,
                                          int(dtype == torch.float64), _ptr(f), _ptr(out)),
                   "lbm_convert")
        return out

    def aos_view(self, f: torch.Tensor) -> torch.Tensor:
        """(*G, Q) view of an f32/f64 state, no copy."""
        if self.storage not in ("f32", "f64"):
            raise ValueError("aos_view needs f32 or f64 storage (deviation storage holds f - w)")
        return f.permute(*range(1, self.D + 1), 0)

    def conservation(self, f: torch.Tensor) -> torch.Tensor:
        """FP64 [sum f, sum f e_0, sum f e_1, (sum f e_2)] as a device tensor."""
        self._check_state(f)
        d = self.desc()
        n = _lib.lib().lbm_reduce_workspace_size(ctypes.byref(d))
        work = torch.empty(n, dtype=torch.uint8, device=self.device)
        out = torch.empty(4, dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().lbm_conservation(_stream(self.device), ctypes.byref(d), _ptr(f),
                                               _ptr(work), _ptr(out)), "lbm_conservation")
        return out[: 1 + self.D]
