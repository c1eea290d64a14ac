"""Drop-in ``LBMPipeNode`` backed by the HIP pipe kernels (lbm_pipe_step).

Mirrors ``src/maddening/nodes/lbm_pipe.py:393-879``: constructor
arguments, validation messages, the pipe, propeller and fill masks,
single-phase and multiphase initial states, the ``propeller_force``
boundary input, and the state keys ``f``, ``density``, ``velocity``,
``tracer`` and ``tracer_f``. ``state["f"]`` (nx, ny, nz, 19) and
``state["tracer_f"]`` (nx, ny, nz, 7) are strided views of the SoA buffers
the kernels use.

The masks are built on the host in float32, as the reference builds them
with jnp float32 arithmetic, so both sides classify every cell the same
way. ``run(state, boundary_inputs, n)`` advances n steps in one launch
sequence.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib
from .lbm import _ptr, _stream
from .node import SimulationNode

_E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
               [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1],
               [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
_W7 = np.array([1 / 4] + [1 / 8] * 6)


def _feq(rho: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """(nx, ny, nz, 19) equilibrium (lbm_pipe.py:144-167) in rho's dtype."""
    e = torch.as_tensor(_E, dtype=u.dtype, device=u.device)
    w = torch.as_tensor(_W, dtype=u.dtype, device=u.device)
    eu = u @ e.T
    usq = (u * u).sum(-1, keepdim=True)
    return w * rho[..., None] * (1.0 + eu * 3.0 + eu * eu * 4.5 - usq * 1.5)


def _teq(c: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """(nx, ny, nz, 7) tracer equilibrium w7 C (1 + e.u/cs2) (lbm_pipe.py:256-271)."""
    e = torch.as_tensor(_E[:7], dtype=u.dtype, device=u.device)
    w = torch.as_tensor(_W7, dtype=u.dtype, device=u.device)
    return w * c[..., None] * (1.0 + (u @ e.T) * 3.0)


class LBMPipeNode(SimulationNode):
    """3-D lattice-Boltzmann fluid in a cylindrical pipe with an actuator-disc
    propeller (same constructor as lbm_pipe.py:493-579, plus keyword-only
    ``storage`` ("f32" or "f64") and ``device``)."""

    def __init__(self, name: str, timestep: float, nx: int = 64, ny: int = 32, nz: int = 32,
                 tau: float = 0.8, pipe_radius: float = 0.9, propeller_x: int = 10,
                 propeller_radius: float = 0.8, propeller_strength: float = 0.0005,
                 initial_velocity: float = 0.0, gravity: float = 0.0,
                 fill_fraction: float = 1.0, tau_tracer: float = 0.6, G: float = 0.0,
                 rho_liquid: float = 1.0, rho_gas: float = 0.25, rho_0: float = 1.0,
                 rho_wall: Optional[float] = None, *, storage: str = "f32", device="cuda"):
        if tau <= 0.5:
            raise ValueError(f"tau must be > 0.5 for stability (got {tau}). "
                             f"nu = (tau - 0.5) / 3 = {(tau - 0.5) / 3:.4f}")
        if tau_tracer <= 0.5:
            raise ValueError(f"tau_tracer must be > 0.5 for stability (got {tau_tracer}).")
        if not 0.0 < fill_fraction <= 1.0:
            raise ValueError(f"fill_fraction must be in (0, 1] (got {fill_fraction}).")
        if G != 0.0:
            if rho_liquid <= rho_gas:
                raise ValueError(f"rho_liquid must be > rho_gas (got {rho_liquid} <= {rho_gas}).")
            if rho_gas <= 0.0:
                raise ValueError(f"rho_gas must be > 0 (got {rho_gas}).")
            if rho_0 <= 0.0:
                raise ValueError(f"rho_0 must be > 0 (got {rho_0}).")
        if storage not in ("f32", "f64"):
            raise ValueError("LBMPipeNode stores f itself: storage must be 'f32' or 'f64'")
        if rho_wall is None:
            rho_wall = rho_liquid
        super().__init__(name, timestep, nx=nx, ny=ny, nz=nz, tau=tau, pipe_radius=pipe_radius,
                         propeller_x=propeller_x, propeller_radius=propeller_radius,
                         propeller_strength=propeller_strength,
                         initial_velocity=initial_velocity, gravity=gravity,
                         fill_fraction=fill_fraction, tau_tracer=tau_tracer, G=G,
                         rho_liquid=rho_liquid, rho_gas=rho_gas, rho_0=rho_0, rho_wall=rho_wall)
        self._nx, self._ny, self._nz = int(nx), int(ny), int(nz)
        self._tau, self._tau_tracer = float(tau), float(tau_tracer)
        self._gravity = float(gravity)
        self._fill_fraction = float(fill_fraction)
        self._G = float(G)
        self._rho_liquid, self._rho_gas = float(rho_liquid), float(rho_gas)
        self._rho_0, self._rho_wall = float(rho_0), float(rho_wall)
        self._wall_mask = self._build_wall_mask(nx, ny, nz, pipe_radius)
        self._propeller_mask = self._build_propeller_mask(nx, ny, nz, propeller_x,
                                                          propeller_radius, pipe_radius)
        self._storage = storage
        self._dtype = torch.float64 if storage == "f64" else torch.float32
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self._device = dev
        flags = self._wall_mask.astype(np.uint8) | (self._propeller_mask.astype(np.uint8) << 1)
        self._flags = torch.as_tensor(np.ascontiguousarray(flags), device=dev)
        self._work = None

    # -- geometry (lbm_pipe.py:581-649) --------------------------------------
    @staticmethod
    def _cross_section_radius(ny, nz):
        y = np.arange(ny, dtype=np.float32)
        z = np.arange(nz, dtype=np.float32)
        yy, zz = np.meshgrid(y, z, indexing="ij")
        cy, cz = (ny - 1) / 2.0, (nz - 1) / 2.0
        return yy, zz, np.sqrt((yy - cy) ** 2 + (zz - cz) ** 2)

    @staticmethod
    def _build_wall_mask(nx, ny, nz, pipe_radius):
        """True at solid wall cells: radius >= pipe_radius * min(ny, nz) / 2."""
        _, _, r = LBMPipeNode._cross_section_radius(ny, nz)
        wall = r >= np.float32(pipe_radius * min(ny, nz) / 2.0)
        return np.ascontiguousarray(np.broadcast_to(wall[None], (nx, ny, nz)))

    @staticmethod
    def _build_propeller_mask(nx, ny, nz, prop_x, prop_radius, pipe_radius):
        """True on the actuator disc at x = prop_x."""
        _, _, r = LBMPipeNode._cross_section_radius(ny, nz)
        mask = np.zeros((nx, ny, nz), dtype=bool)
        # jnp's .at[prop_x].set drops an out-of-range index: no disc then
        if -nx <= prop_x < nx:
            mask[prop_x] = r < np.float32(prop_radius * pipe_radius * min(ny, nz) / 2.0)
        return mask

    def _compute_fill_mask(self):
        """Initially liquid cells: below the fill_fraction percentile of the
        interior z-coordinates, pulled in near the wall at the surface."""
        ny, nz = self._ny, self._nz
        pipe_radius = self.params["pipe_radius"]
        cy, cz = (ny - 1) / 2.0, (nz - 1) / 2.0
        max_r = pipe_radius * min(ny, nz) / 2.0
        yy, zz, _ = self._cross_section_radius(ny, nz)
        interior = ~self._wall_mask[0]
        z_cutoff = float(np.percentile(zz[interior], self._fill_fraction * 100))
        r = np.sqrt((yy - cy) ** 2 + (zz - cz) ** 2)
        near = ((max_r - r) < 1.5) & (np.abs(zz - z_cutoff) < 2.0)
        cross = (zz <= z_cutoff) & interior & ~near
        return np.ascontiguousarray(np.broadcast_to(cross[None], (self._nx, ny, nz)))

    @property
    def viscosity(self) -> float:
        return (self.params["tau"] - 0.5) / 3.0

    def halo_width(self) -> dict:
        return {0: 1, 1: 1, 2: 1}

    # -- state ----------------------------------------------------------------------
    def _soa_view(self, soa: torch.Tensor) -> torch.Tensor:
        return soa.permute(1, 2, 3, 0)

    def _soa(self, aos: torch.Tensor, q: int) -> torch.Tensor:
        shape = (self._nx, self._ny, self._nz, q)
        if tuple(aos.shape) != shape:
            raise ValueError(f"expected a state of shape {shape}")
        t = aos.to(device=self._device, dtype=self._dtype).permute(3, 0, 1, 2)
        return t if t.is_contiguous() else t.contiguous()

    def initial_state(self) -> dict:
        """Single phase: rho = 1, uniform x-velocity in the pipe, tracer from the
        fill mask. Multiphase: tanh liquid/gas density profile at the fill
        level, tracer from density (lbm_pipe.py:664-749)."""
        nx, ny, nz = self._nx, self._ny, self._nz
        dev, dt = self._device, self._dtype
        wall = torch.as_tensor(self._wall_mask, device=dev)
        fluid = ~wall
        u0 = self.params["initial_velocity"]
        vel = torch.zeros((nx, ny, nz, 3), dtype=torch.float32, device=dev)
        vel[..., 0] = torch.where(fluid, torch.tensor(float(u0), device=dev), torch.tensor(0.0, device=dev))
        if self._G != 0.0:
            interior = ~self._wall_mask[0]
            zz = self._cross_section_radius(ny, nz)[1]
            z_cut = float(nz) if self._fill_fraction >= 1.0 else float(
                np.percentile(zz[interior], self._fill_fraction * 100))
            zc = np.arange(nz, dtype=np.float32)
            phase = (0.5 * (1.0 + np.tanh((np.float32(z_cut) - zc) / np.float32(5.0)))).astype(
                np.float32)
            dens = (self._rho_gas + (self._rho_liquid - self._rho_gas)
                    * np.broadcast_to(phase[None, None, :], (nx, ny, nz))).astype(np.float32)
            dens = np.where(self._wall_mask, np.float32(self._rho_gas), dens).astype(np.float32)
            density = torch.as_tensor(dens, device=dev)
            tr = np.clip((dens - self._rho_gas) / (self._rho_liquid - self._rho_gas), 0.0, 1.0)
            tracer = torch.as_tensor(np.where(self._wall_mask, 0.0, tr).astype(np.float32),
                                     device=dev)
        else:
            density = torch.ones((nx, ny, nz), dtype=torch.float32, device=dev)
            liquid = fluid if self._fill_fraction >= 1.0 else torch.as_tensor(
                self._compute_fill_mask(), device=dev)
            tracer = liquid.to(torch.float32)
        f = _feq(density.to(dt), vel.to(dt))
        g = _teq(tracer.to(dt), vel.to(dt))
        return {"f": self._soa_view(f.permute(3, 0, 1, 2).contiguous()),
                "density": density.to(dt), "velocity": vel.to(dt), "tracer": tracer.to(dt),
                "tracer_f": self._soa_view(g.permute(3, 0, 1, 2).contiguous())}

    # -- step -----------------------------------------------------------------------
    def _desc(self, strength: float) -> _lib.LbmPipeDesc:
        d = _lib.LbmPipeDesc()
        d.dims[0], d.dims[1], d.dims[2] = self._nx, self._ny, self._nz
        d.storage = _lib.STORAGE_NAMES[self._storage]
        d.multiphase = int(self._G != 0.0)
        d.tau, d.tau_tracer = self._tau, self._tau_tracer
        d.propeller_strength = float(strength)
        d.gravity = self._gravity
        d.G, d.rho_0, d.rho_wall = self._G, self._rho_0, self._rho_wall
        d.rho_liquid, d.rho_gas = self._rho_liquid, self._rho_gas
        return d

    def update(self, state: dict, boundary_inputs: dict, dt: float) -> dict:
        """One LBMPipeNode.update (lbm_pipe.py:751-879)."""
        return self.run(state, boundary_inputs, 1)

    def run(self, state: dict, boundary_inputs: dict, nsteps: int) -> dict:
        """``nsteps`` updates with the same boundary inputs in one launch sequence."""
        if self._device.type != "cuda":
            raise RuntimeError("LBMPipeNode steps on the GPU: construct it with a cuda device")
        strength = boundary_inputs.get("propeller_force", self.params["propeller_strength"])
        strength = float(strength.item() if torch.is_tensor(strength) else strength)
        d = self._desc(strength)
        L = _lib.lib()
        nbytes = L.lbm_pipe_workspace_size(ctypes.byref(d))
        if nbytes == 0:
            _lib.check(-1, "lbm_pipe_workspace_size")
        if self._work is None or self._work.numel() < nbytes:
            self._work = torch.empty(nbytes, dtype=torch.uint8, device=self._device)
        f = self._soa(state["f"], 19)
        g = self._soa(state["tracer_f"], 7)
        shape = (self._nx, self._ny, self._nz)
        f_out = torch.empty_like(f)
        g_out = torch.empty_like(g)
        rho = torch.empty(shape, dtype=self._dtype, device=self._device)
        u = torch.empty(shape + (3,), dtype=self._dtype, device=self._device)
        tracer = torch.empty(shape, dtype=self._dtype, device=self._device)
        _lib.check(L.lbm_pipe_step(_stream(self._device), ctypes.byref(d), _ptr(f), _ptr(f_out),
                                   _ptr(g), _ptr(g_out), _ptr(self._flags), _ptr(self._work),
                                   _ptr(rho), _ptr(u), _ptr(tracer), int(nsteps)),
                   "lbm_pipe_step")
        return {"f": self._soa_view(f_out), "density": rho, "velocity": u, "tracer": tracer,
                "tracer_f": self._soa_view(g_out)}
