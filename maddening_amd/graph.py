"""Minimal graph harness: the calling contract MADDENING's GraphManager gives
a node, for running the LBM node on the GPU without the reference.

It mirrors ``src/maddening/core/graph_manager.py``:

* ``add_node`` (:1578-1591) stores ``node.update`` and the node's initial
  state.
* ``add_edge`` (:1593-1627) and ``add_external_input`` (:1629-1655) route
  values into ``boundary_inputs``.
* ``compile`` (:2006) fixes the schedule. ``step`` (:2638) and ``run``
  (:2663) advance one base step per call.
* ``run_scan`` / ``run_scan_with_history`` (:2697-2805) replace ``lax.scan``.
  ``run_sweep`` (:2807-2881) replaces ``vmap(scan)`` over a leading batch
  axis of the initial states.

Boundary inputs resolve as in ``_build_step_fn`` (:2387-2449). A forward
edge reads the source's state from this step, and a back edge (one that
closes a cycle) reads the previous step's state. A field missing from the
source's state is looked up in the source's ``compute_boundary_fluxes``
output. Edges to the same field are summed when ``additive``; external
inputs are applied last.

Execution maps onto the HIP path rather than onto XLA:

* ``run_scan`` of a graph without edges gives each node's whole scan to its
  ``run(state, boundary_inputs, nsteps)`` when the node has one. For the
  LBM node that is one launch sequence: nsteps + 1 kernels, no host work
  per step.
* ``run_sweep`` of a single f32 LBM node without edges advances every batch
  member in one batched launch per step (``LatticeBatch``). Other graphs
  loop over the members.

Coupling groups, multirate, adaptive stepping, sharding and observers'
event payloads are outside the LBM path and are not provided.
"""
from __future__ import annotations

from collections import defaultdict
from dataclasses import dataclass
from typing import Any, Callable, Optional

import torch


@dataclass(frozen=True)
class EdgeSpec:
    """A data dependency (graph_manager.py EdgeSpec)."""
    source_node: str
    target_node: str
    source_field: str
    target_field: str
    transform: Optional[Callable] = None
    additive: bool = False


@dataclass(frozen=True)
class ExternalInputSpec:
    """A value injected from outside each step (graph_manager.py ExternalInputSpec)."""
    target_node: str
    target_field: str
    shape: tuple = ()
    dtype: Any = torch.float32


@dataclass
class _NodeSpec:
    node: Any
    update_fn: Callable
    timestep: float


def _overrides_fluxes(node) -> bool:
    base = getattr(type(node), "__mro__", ())
    impl = type(node).compute_boundary_fluxes
    for cls in base[1:]:
        if "compute_boundary_fluxes" in cls.__dict__:
            return impl is not cls.__dict__["compute_boundary_fluxes"]
    return False


class GraphManager:
    """Nodes, edges and external inputs advanced together (see module docstring)."""

    def __init__(self) -> None:
        self._nodes: dict[str, _NodeSpec] = {}
        self._edges: list[EdgeSpec] = []
        self._external_inputs: list[ExternalInputSpec] = []
        self._state: dict[str, dict] = {}
        self._schedule: list[str] = []
        self._back_edges: set[EdgeSpec] = set()
        self._compiled = False
        self._observers: list[Callable] = []

    # -- construction ---------------------------------------------------------
    def add_node(self, node) -> None:
        """Register a node and initialise its state."""
        if node.name in self._nodes:
            raise ValueError(f"Node '{node.name}' already exists in the graph.")
        self._nodes[node.name] = _NodeSpec(node=node, update_fn=node.update,
                                           timestep=node.delta_t)
        self._state[node.name] = node.initial_state()
        self._compiled = False

    def add_edge(self, source: str, target: str, source_field: str, target_field: str,
                 transform: Optional[Callable] = None, additive: bool = False) -> None:
        """Add a data-dependency edge between two nodes."""
        self._edges.append(EdgeSpec(source, target, source_field, target_field, transform,
                                    additive))
        self._compiled = False

    def add_external_input(self, target_node: str, target_field: str, shape: tuple = (),
                           dtype: Any = torch.float32) -> None:
        """Declare an input supplied through ``external_inputs`` every step."""
        self._external_inputs.append(ExternalInputSpec(target_node, target_field, tuple(shape),
                                                       dtype))
        self._compiled = False

    def remove_node(self, name: str) -> None:
        """Remove a node and the edges and external inputs that reference it."""
        if name not in self._nodes:
            raise KeyError(f"No node named '{name}'.")
        del self._nodes[name]
        self._state.pop(name, None)
        self._edges = [e for e in self._edges if name not in (e.source_node, e.target_node)]
        self._external_inputs = [e for e in self._external_inputs if e.target_node != name]
        self._compiled = False

    def validate(self) -> list[str]:
        """Problems that would make ``compile`` fail (empty when the graph is sound)."""
        errs = []
        for e in self._edges:
            for end in (e.source_node, e.target_node):
                if end not in self._nodes:
                    errs.append(f"Edge {e.source_node}.{e.source_field} -> "
                                f"{e.target_node}.{e.target_field} references unknown node "
                                f"'{end}'.")
        for ei in self._external_inputs:
            if ei.target_node not in self._nodes:
                errs.append(f"External input '{ei.target_field}' targets unknown node "
                            f"'{ei.target_node}'.")
        return errs

    def compile(self) -> None:
        """Fix the update order: a topological order of the edges with ties in
        insertion order; edges that close a cycle become back edges."""
        errs = self.validate()
        if errs:
            raise ValueError("; ".join(errs))
        names = list(self._nodes)
        indeg = {n: 0 for n in names}
        out = defaultdict(list)
        for e in self._edges:
            if e.source_node != e.target_node:
                out[e.source_node].append(e.target_node)
                indeg[e.target_node] += 1
        order: list[str] = []
        remaining = set(names)
        while remaining:
            ready = [n for n in names if n in remaining and indeg[n] == 0]
            # a cycle: break it at the earliest-inserted node left
            pick = ready[0] if ready else next(n for n in names if n in remaining)
            order.append(pick)
            remaining.discard(pick)
            for t in out[pick]:
                indeg[t] -= 1
        pos = {n: i for i, n in enumerate(order)}
        self._schedule = order
        self._back_edges = {e for e in self._edges if pos[e.source_node] >= pos[e.target_node]}
        self._compiled = True

    def _ensure_compiled(self) -> None:
        if not self._compiled:
            self.compile()

    # -- state -------------------------------------------------------------------
    def get_node_state(self, name: str) -> dict:
        return self._state[name]

    def set_node_state(self, name: str, state: dict) -> None:
        if name not in self._nodes:
            raise KeyError(f"No node named '{name}'.")
        self._state[name] = dict(state)

    @property
    def state(self) -> dict:
        return self._state

    def add_observer(self, callback: Callable) -> None:
        """``callback(event, data)`` after every ``step``/``run`` step."""
        self._observers.append(callback)

    def _notify(self, event: str, data) -> None:
        for cb in self._observers:
            cb(event, data)

    # -- one step ----------------------------------------------------------------
    def _step_fn(self, full_state: dict, external_inputs: dict) -> dict:
        new_state = dict(full_state)
        fluxes: dict[str, dict] = {}
        by_target = defaultdict(list)
        for e in self._edges:
            by_target[e.target_node].append(e)
        for name in self._schedule:
            bi: dict[str, Any] = {}
            for e in by_target[name]:
                src_all = full_state if e in self._back_edges else new_state
                src = src_all.get(e.source_node, {})
                if e.source_field in src:
                    value = src[e.source_field]
                else:
                    value = fluxes[e.source_node][e.source_field]
                if e.transform is not None:
                    value = e.transform(value)
                if e.additive and e.target_field in bi:
                    bi[e.target_field] = bi[e.target_field] + value
                else:
                    bi[e.target_field] = value
            for k, v in external_inputs.get(name, {}).items():
                bi[k] = v
            spec = self._nodes[name]
            new_state[name] = spec.update_fn(new_state[name], bi, spec.timestep)
            if _overrides_fluxes(spec.node):
                fl = spec.node.compute_boundary_fluxes(new_state[name], bi, spec.timestep)
                if fl:
                    fluxes[name] = fl
        return new_state

    def _defaults(self, external_inputs: Optional[dict]) -> dict:
        if external_inputs is not None:
            return external_inputs
        ext: dict[str, dict] = {}
        for ei in self._external_inputs:
            node = self._nodes[ei.target_node].node
            dev = getattr(node, "_device", None)
            ext.setdefault(ei.target_node, {})[ei.target_field] = torch.zeros(
                ei.shape, dtype=ei.dtype, device=dev)
        return ext

    def step(self, external_inputs: Optional[dict] = None) -> dict:
        """Advance every node by one base step; returns the new state."""
        self._ensure_compiled()
        self._state = self._step_fn(self._state, self._defaults(external_inputs))
        self._notify("step", self._state)
        return self._state

    def run(self, n_steps: int, callback: Optional[Callable] = None,
            external_inputs: Optional[dict] = None) -> None:
        """``n_steps`` calls of ``step`` with ``callback(i, state)`` after each."""
        self._ensure_compiled()
        ext = self._defaults(external_inputs)
        for i in range(n_steps):
            self._state = self._step_fn(self._state, ext)
            self._notify("step", self._state)
            if callback is not None:
                callback(i, self._state)

    # -- scans ---------------------------------------------------------------------
    def _fusable(self) -> bool:
        return not self._edges and all(hasattr(s.node, "run") for s in self._nodes.values())

    def _scan(self, state: dict, n_steps: int, ext: dict) -> dict:
        if n_steps <= 0:
            return state
        if self._fusable():
            return {name: self._nodes[name].node.run(state[name], ext.get(name, {}), n_steps)
                    for name in self._schedule}
        for _ in range(n_steps):
            state = self._step_fn(state, ext)
        return state

    def run_scan(self, n_steps: int, external_inputs: Optional[dict] = None) -> dict:
        """``n_steps`` steps with static external inputs; returns the final state."""
        self._ensure_compiled()
        self._state = self._scan(self._state, n_steps, self._defaults(external_inputs))
        return self._state

    def run_scan_with_history(self, n_steps: int, external_inputs: Optional[dict] = None):
        """Like ``run_scan`` and also every step's state, stacked on a new
        leading axis of length ``n_steps`` (the state after each step)."""
        self._ensure_compiled()
        ext = self._defaults(external_inputs)
        hist: list[dict] = []
        state = self._state
        for _ in range(n_steps):
            state = self._step_fn(state, ext)
            hist.append(state)
        self._state = state
        return state, _stack(hist)

    def run_sweep(self, n_steps: int, initial_states: dict, external_inputs: Optional[dict] = None,
                  return_history: bool = False):
        """Run a batch of simulations whose initial states carry a leading batch
        axis on every leaf; returns the batched final states (and, with
        ``return_history``, histories shaped (batch, n_steps, ...)). The
        graph's own state is left unchanged."""
        self._ensure_compiled()
        ext = self._defaults(external_inputs)
        sizes = {int(v.shape[0]) for st in initial_states.values() for v in st.values()}
        if len(sizes) != 1:
            raise ValueError("every leaf of initial_states needs the same leading batch size")
        batch = sizes.pop()
        if not return_history:
            fast = _batched_lbm_sweep(self, n_steps, initial_states, ext, batch)
            if fast is not None:
                return fast
        finals, hists = [], []
        saved = self._state
        try:
            for b in range(batch):
                st = {n: {k: v[b] for k, v in s.items()} for n, s in initial_states.items()}
                full = dict(saved)
                full.update(st)
                if return_history:
                    h = []
                    for _ in range(n_steps):
                        full = self._step_fn(full, ext)
                        h.append({n: full[n] for n in initial_states})
                    hists.append(_stack(h))
                else:
                    full = self._scan(full, n_steps, ext)
                finals.append({n: full[n] for n in initial_states})
        finally:
            self._state = saved
        out = _stack(finals)
        return (out, _stack(hists)) if return_history else out


def _stack(states: list[dict]) -> dict:
    if not states:
        return {}
    return {n: {k: torch.stack([torch.as_tensor(s[n][k]) for s in states])
                for k in states[0][n]} for n in states[0]}


def _batched_lbm_sweep(gm: GraphManager, n_steps: int, initial_states: dict, ext: dict,
                       batch: int):
    """The sweep of one f32 LBMNode without edges, uniform (or no) body force
    and no pressure faces, as one batched launch per step. Returns None when
    the graph is not of that form."""
    from .batch import LatticeBatch
    from .node import LBMNode
    if gm._edges or len(gm._nodes) != 1 or n_steps < 1:
        return None
    name, spec = next(iter(gm._nodes.items()))
    node = spec.node
    if not isinstance(node, LBMNode) or node.kernel.storage != "f32" or name not in initial_states:
        return None
    if node.kernel.bb == "halfway":
        return None
    bi = ext.get(name, {})
    if "inlet_pressure" in bi or "outlet_pressure" in bi:
        return None
    force = bi.get("body_force")
    if force is not None:
        fl = torch.as_tensor(force).flatten().tolist() if torch.is_tensor(force) else list(force)
        if torch.as_tensor(force).ndim != 1:
            return None
        force = [float(v) for v in fl]
    init = initial_states[name]
    k = node.kernel
    masks = []
    for b in range(batch):
        if "wall_mask_update" in bi:
            m = torch.as_tensor(bi["wall_mask_update"])
        elif "wall_mask" in init:
            m = init["wall_mask"][b]
        else:
            m = node._wall_mask
        masks.append(torch.as_tensor(m).to(torch.uint8).cpu().numpy())
    specs = [{"lattice": k.lat.name, "shape": k.grid_shape, "tau": k.tau, "bb": "reference",
              "mask": masks[b], **({"force": force} if force else {})} for b in range(batch)]
    lb = LatticeBatch(specs, device=k.device)
    try:
        lb.set_states([node._soa(init["f"][b]) for b in range(batch)])
        lb.run(n_steps)
        finals = lb.states(n_steps)
        out = {"f": [], "density": [], "velocity": [], "pressure": []}
        for b, (kb, fb) in enumerate(zip(lb.kernels, finals)):
            rho, u, p = kb.moments(fb, mask=lb.masks[b], force=force)
            out["f"].append(kb.aos_view(fb))
            out["density"].append(rho)
            out["velocity"].append(u)
            out["pressure"].append(p)
        res = {key: torch.stack(v) for key, v in out.items()}
        if "wall_mask" in init:
            res["wall_mask"] = torch.as_tensor(init["wall_mask"])
        return {name: res}
    finally:
        lb.close()
