"""Drop-in ``LBMNode`` for MADDENING's node contract, backed by the HIP kernels.

Mirrors ``src/maddening/nodes/lbm.py:531-1036`` (constructor arguments,
validation messages, state keys, boundary inputs, fluxes) and the
``SimulationNode`` contract of ``src/maddening/core/node.py:70-480``. State
tensors live on the GPU; ``state["f"]`` has the reference's shape
``(*grid_shape, Q)`` and is a strided view of the SoA buffer the kernels
use, so no transpose happens per step.

When MADDENING itself is importable the node subclasses its
``SimulationNode``; otherwise a local ABC with the same methods is used.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np
import torch

from .lattice import LatticeDescriptor, d2q9, d3q19
from .lbm import LBMKernel

try:  # pragma: no cover - only where the reference package is installed
    from maddening.core.node import BoundaryInputSpec, SimulationNode  # type: ignore
except Exception:  # the reference (and JAX) are optional

    @dataclass(frozen=True)
    class BoundaryInputSpec:
        """Expected boundary input (core/node.py:23-46)."""
        shape: tuple = ()
        dtype: Any = None
        default: Any = None
        coupling_type: str = "replacive"
        description: str = ""
        expected_units: Optional[str] = None

    class SimulationNode(ABC):
        """Node contract (core/node.py:70-480): a descriptor exposing
        ``initial_state()`` and a pure ``update(state, boundary_inputs, dt)``."""

        meta = None

        def __init__(self, name: str, timestep: float, **params):
            self.name = name
            self.delta_t = float(timestep)
            self.geometry_source = params.pop("geometry_source", None)
            self.params = dict(params)

        @abstractmethod
        def initial_state(self) -> dict:
            ...

        @abstractmethod
        def update(self, state: dict, boundary_inputs: dict, dt: float) -> dict:
            ...

        def halo_width(self) -> dict:
            return {}

        def boundary_input_spec(self) -> dict:
            return {}

        def compute_boundary_fluxes(self, state: dict, boundary_inputs: dict,
                                    dt: float) -> dict:
            return {}

        def derivatives(self, state: dict, boundary_inputs: dict) -> dict:
            raise NotImplementedError

        def to_dict(self) -> dict:
            return {"type": type(self).__name__, "name": self.name,
                    "timestep": self.delta_t, "params": dict(self.params)}


FACE_MAP = {
    "x_min": (0, "min"), "x_max": (0, "max"),
    "y_min": (1, "min"), "y_max": (1, "max"),
    "z_min": (2, "min"), "z_max": (2, "max"),
}


class LBMNode(SimulationNode):
    """General 3-D/2-D lattice-Boltzmann node (BGK + Guo forcing, wall
    bounce-back) on an arbitrary wall mask, computed by the gfx950 kernels.

    Same constructor as the reference (lbm.py:604-615) plus keyword-only
    build options: ``storage`` ("f32" or "f64"), ``bb`` ("reference", the
    reference's wall rule, or "halfway"), ``device``, ``scheme`` and
    ``pair`` (two steps per launch, see ``LBMKernel``; with "auto" each
    D3Q19 ``update`` is one launch).
    """

    def __init__(self, name: str, timestep: float, grid_shape: tuple = (64, 32, 32),
                 viscosity: float = 0.1, lattice: str = "D3Q19",
                 wall_mask: Optional[np.ndarray] = None, inlet_face: str = "x_min",
                 outlet_face: str = "x_max", geometry_source: Optional[str] = None, *,
                 storage: str = "f32", bb: str = "reference", device="cuda",
                 scheme: str = "auto", pair: str = "auto"):
        if lattice.upper() == "D3Q19":
            lat = d3q19()
        elif lattice.upper() == "D2Q9":
            lat = d2q9()
        else:
            raise ValueError(f"Unknown lattice '{lattice}'. Supported: 'D3Q19', 'D2Q9'.")
        if len(grid_shape) != lat.D:
            raise ValueError(f"grid_shape has {len(grid_shape)} dims but lattice {lat.name} "
                             f"requires {lat.D} dims.")
        tau = 0.5 + viscosity / lat.cs2
        if tau <= 0.5:
            raise ValueError(f"tau must be > 0.5 (got {tau}). Increase viscosity.")
        for face_name, face_label in (("inlet_face", inlet_face), ("outlet_face", outlet_face)):
            if face_label not in FACE_MAP:
                raise ValueError(f"Unknown {face_name} '{face_label}'. "
                                 f"Supported: {list(FACE_MAP.keys())}")
            if FACE_MAP[face_label][0] >= lat.D:
                raise ValueError(f"{face_name} '{face_label}' uses axis {FACE_MAP[face_label][0]} "
                                 f"but lattice {lat.name} only has {lat.D} dimensions.")
        if storage not in ("f32", "f64"):
            raise ValueError("LBMNode state holds f itself: storage must be 'f32' or 'f64'")
        if wall_mask is not None and tuple(np.shape(wall_mask)) != tuple(grid_shape):
            raise ValueError(f"wall_mask shape {tuple(np.shape(wall_mask))} != grid_shape "
                             f"{tuple(grid_shape)}")
        super().__init__(name, timestep, grid_shape=grid_shape, viscosity=viscosity,
                         lattice=lattice, inlet_face=inlet_face, outlet_face=outlet_face,
                         geometry_source=geometry_source)
        self._lat = lat
        self._grid_shape = tuple(int(n) for n in grid_shape)
        self._D, self._Q = lat.D, lat.Q
        self._tau = tau
        self._cs2 = lat.cs2
        self._inlet_face = inlet_face
        self._outlet_face = outlet_face
        self._kernel = LBMKernel(self._grid_shape, lattice=lat.name, tau=tau, storage=storage,
                                 bb=bb, scheme=scheme, pair=pair, device=device)
        self._bb = bb
        self._device = self._kernel.device
        self._dtype = self._kernel.dtype
        if wall_mask is not None:
            wm = np.asarray(wall_mask)
            self._has_walls = bool(np.any(wm))
            self._wall_mask = torch.as_tensor(wm.astype(np.uint8), device=self._device)
        else:
            self._has_walls = False
            self._wall_mask = torch.zeros(self._grid_shape, dtype=torch.uint8, device=self._device)
        # a mask without walls runs the periodic kernels: the wall rule then
        # changes no cell, so the values are the same, and the periodic
        # two-step form is the leanest
        self._open_kernel = None
        self._mask_seen = None  # (mask tensor, (data_ptr, version)) last checked
        self._mask_walls = True

    # -- properties (lbm.py:690-705) ---------------------------------------
    @property
    def tau(self) -> float:
        return self._tau

    @property
    def viscosity(self) -> float:
        return self._cs2 * (self._tau - 0.5)

    @property
    def lattice(self) -> LatticeDescriptor:
        return self._lat

    @property
    def kernel(self) -> LBMKernel:
        return self._kernel

    def halo_width(self) -> dict:
        """One ghost cell per axis (lbm.py:709-717)."""
        return {axis: 1 for axis in range(self._D)}

    # -- state -------------------------------------------------------------
    def initial_state(self) -> dict:
        """rho = 1, u = 0, f = feq, p = cs2, wall_mask uint8 (lbm.py:719-745)."""
        k = self._kernel
        f = k.aos_view(k.equilibrium())
        density = torch.ones(self._grid_shape, dtype=self._dtype, device=self._device)
        return {
            "f": f,
            "density": density,
            "velocity": torch.zeros(self._grid_shape + (self._D,), dtype=self._dtype,
                                    device=self._device),
            "pressure": density * self._cs2,
            "wall_mask": self._wall_mask.clone(),
        }

    def _soa(self, f: torch.Tensor) -> torch.Tensor:
        """SoA (Q, *G) tensor for a (*G, Q) state, a view when it already is one."""
        if tuple(f.shape) != self._grid_shape + (self._Q,):
            raise ValueError(f"state['f'] must have shape {self._grid_shape + (self._Q,)}")
        soa = f.to(device=self._device, dtype=self._dtype).permute(self._D, *range(self._D))
        return soa if soa.is_contiguous() else soa.contiguous()

    def _inputs(self, state: dict, boundary_inputs: dict):
        # wall precedence: wall_mask_update > state["wall_mask"] > ctor mask (lbm.py:914-919)
        if "wall_mask_update" in boundary_inputs:
            mask = torch.as_tensor(boundary_inputs["wall_mask_update"], device=self._device)
        elif "wall_mask" in state:
            mask = torch.as_tensor(state["wall_mask"], device=self._device)
        else:
            mask = self._wall_mask
        force = boundary_inputs.get("body_force", None)
        if force is not None:
            # the reference casts the force to float32 (lbm.py:924-931); a
            # uniform force given on the host goes to the kernels as numbers
            # without a trip through device memory
            if torch.is_tensor(force):
                force = force.to(device=self._device, dtype=torch.float32)
                if force.ndim == 1:
                    force = force.tolist()
            else:
                arr = np.asarray(force, dtype=np.float32)
                force = arr.tolist() if arr.ndim == 1 else torch.as_tensor(arr, device=self._device)
            if isinstance(force, list):
                if len(force) != self._D:
                    raise ValueError(f"uniform body_force must have shape ({self._D},)")
                force = [float(v) for v in force]
        # Zou-He faces: rho_p = P / cs2 at the configured faces (lbm.py:950-968)
        faces = None
        p_in = boundary_inputs.get("inlet_pressure")
        p_out = boundary_inputs.get("outlet_pressure")
        if p_in is not None or p_out is not None:
            def num(v):
                return float(v.detach()) if torch.is_tensor(v) else float(v)
            faces = (None if p_in is None else (self._inlet_face, num(p_in) / self._cs2),
                     None if p_out is None else (self._outlet_face, num(p_out) / self._cs2))
        return mask, force, faces

    def _kernel_for(self, mask: torch.Tensor):
        """(kernel, mask argument) for a call on ``mask``. Whether the mask
        has a wall is read from the device once per mask tensor and version
        (one synchronisation when the mask changes, none per step)."""
        key = (mask.data_ptr(), mask._version)
        if self._mask_seen is None or self._mask_seen[0] is not mask or self._mask_seen[1] != key:
            self._mask_walls = bool(mask.any())
            self._mask_seen = (mask, key)
        if self._mask_walls:
            return self._kernel, mask.to(torch.uint8)
        if self._open_kernel is None:
            k = self._kernel
            self._open_kernel = LBMKernel(self._grid_shape, lattice=self._lat.name, tau=self._tau,
                                          storage=k.storage, bb="none", scheme=k.scheme,
                                          pair=k.pair, device=self._device)
        return self._open_kernel, None

    def update(self, state: dict, boundary_inputs: dict, dt: float) -> dict:
        """One LBM step (lbm.py:899-986)."""
        return self.run(state, boundary_inputs, 1)

    def run(self, state: dict, boundary_inputs: dict, nsteps: int) -> dict:
        """``nsteps`` updates with the same boundary inputs in one kernel
        sequence; equal to calling ``update`` nsteps times.

        Differentiable: when ``state["f"]``, a tensor ``body_force`` or a
        tensor ``inlet_pressure``/``outlet_pressure`` requires grad (and grad
        mode is on), the step runs as an autograd rollout whose backward is
        the HIP adjoint (every state kept when they fit in free device
        memory, sqrt(N) checkpointing otherwise). That is the counterpart of
        jax.grad through the reference's update."""
        mask, force, faces = self._inputs(state, boundary_inputs)
        k, kmask = self._kernel_for(mask)
        raw_force = boundary_inputs.get("body_force")
        raw_p = (boundary_inputs.get("inlet_pressure"), boundary_inputs.get("outlet_pressure"))
        wants_grad = torch.is_grad_enabled() and (
            (torch.is_tensor(state["f"]) and state["f"].requires_grad)
            or (torch.is_tensor(raw_force) and raw_force.requires_grad)
            or any(torch.is_tensor(v) and v.requires_grad for v in raw_p))
        if wants_grad:
            from .adjoint import rollout
            gfaces = None
            if faces is not None:
                # rho_p = P / cs2, kept differentiable when P is a tensor
                def rho_of(v):
                    return v / self._cs2 if torch.is_tensor(v) else float(v) / self._cs2
                gfaces = (None if raw_p[0] is None else (self._inlet_face, rho_of(raw_p[0])),
                          None if raw_p[1] is None else (self._outlet_face, rho_of(raw_p[1])))
            gforce = None
            if raw_force is not None:
                gforce = raw_force if torch.is_tensor(raw_force) else torch.as_tensor(
                    np.asarray(raw_force, np.float32), device=self._device)
                # float32 values as on the non-differentiable path (and in the
                # reference); the cast passes gradients through unchanged
                gforce = gforce.to(self._device).float()
            f_new, rho, u, p = rollout(k, self._soa(state["f"]), nsteps, force=gforce,
                                       mask=kmask, faces=gfaces)
        else:
            f_new, rho, u, p = k.step(self._soa(state["f"]), nsteps, mask=kmask, force=force,
                                      faces=faces)
        result = {
            "f": k.aos_view(f_new),
            "density": rho.to(self._dtype),
            "velocity": u.to(self._dtype),
            "pressure": p.to(self._dtype),
        }
        if "wall_mask" in state:
            result["wall_mask"] = state["wall_mask"]
        return result

    def derivatives(self, state: dict, boundary_inputs: dict) -> dict:
        raise NotImplementedError(
            "LBMNode uses a discrete lattice Boltzmann update, "
            "not a continuous ODE. derivatives() is not applicable.")

    def compute_boundary_fluxes(self, state: dict, boundary_inputs: dict, dt: float) -> dict:
        """Mean fluid pressure on the outlet face (lbm.py:995-1017)."""
        axis, side = FACE_MAP[self._outlet_face]
        idx = 0 if side == "min" else -1
        p_face = state["pressure"].select(axis, idx)
        wall_face = self._wall_mask.select(axis, idx) != 0
        count = (~wall_face).sum().clamp(min=1)
        return {"outlet_pressure_avg": torch.where(wall_face, 0.0, p_face).sum() / count}

    def boundary_input_spec(self) -> dict:
        """Same keys as lbm.py:1019-1036."""
        return {
            "inlet_pressure": BoundaryInputSpec(shape=(), description="Zou-He pressure at inlet face"),
            "outlet_pressure": BoundaryInputSpec(shape=(), description="Zou-He pressure at outlet face"),
            "body_force": BoundaryInputSpec(shape=(*self._grid_shape, self._D),
                                            coupling_type="additive",
                                            description="External body force field"),
            "wall_mask_update": BoundaryInputSpec(shape=self._grid_shape,
                                                  description="Runtime wall mask override (True=wall)"),
        }
