"""Node-state checkpoints in the reference's on-disk format.

Files written here load with MADDENING's ``load_state`` and vice versa
(``src/maddening/core/simulation/checkpoint.py:69-199``). Each node field is
one ``.npz`` entry named ``"<node>/<field>"``, and ``f`` is the reference's
``(*grid_shape, Q)`` float32 array. The GPU state keeps the SoA buffer, and
conversion happens only here. An optional sidecar
``<file>.npz.manifest.json`` records the schema version, SHA-256 and size,
in the layout of ``checkpoint.py:272-330``.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path
from typing import Mapping, Optional

import numpy as np
import torch

SCHEMA_VERSION = 1
_META_KEY = "_meta"


class CheckpointIntegrityError(ValueError):
    """The manifest does not match the checkpoint (hash, size or schema)."""


def _to_numpy(value) -> np.ndarray:
    if isinstance(value, torch.Tensor):
        return value.detach().contiguous().cpu().numpy()
    return np.asarray(value)


def file_sha256(path) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for chunk in iter(lambda: fh.read(1 << 16), b""):
            h.update(chunk)
    return h.hexdigest()


def save_state(states: Mapping[str, Mapping[str, object]], path, *, manifest: bool = True,
               extra: Optional[dict] = None, meta: Optional[Mapping[str, object]] = None) -> Path:
    """Write ``{node_name: state_dict}`` to ``path`` (``.npz`` appended if
    missing); returns the written path."""
    path = Path(path)
    arrays = {}
    for node, st in states.items():
        if "/" in node:
            raise ValueError(f"node name '{node}' must not contain '/'")
        for field, value in st.items():
            arrays[f"{node}/{field}"] = _to_numpy(value)
    for field, value in (meta or {}).items():
        arrays[f"{_META_KEY}/{field}"] = _to_numpy(value)
    np.savez(path, **arrays)
    out = path if path.suffix == ".npz" else path.with_suffix(path.suffix + ".npz")
    if manifest:
        man = {"schema_version": SCHEMA_VERSION, "sha256": file_sha256(out),
               "size_bytes": out.stat().st_size, "extra": dict(extra or {})}
        out.with_suffix(out.suffix + ".manifest.json").write_text(json.dumps(man, indent=2),
                                                                  encoding="utf-8")
    return out


def verify_manifest(path) -> None:
    """Raise CheckpointIntegrityError when the sidecar manifest disagrees."""
    path = Path(path)
    mpath = path.with_suffix(path.suffix + ".manifest.json")
    if not mpath.exists():
        raise FileNotFoundError(f"No manifest beside checkpoint: {mpath}")
    man = json.loads(mpath.read_text(encoding="utf-8"))
    if man.get("schema_version") != SCHEMA_VERSION:
        raise CheckpointIntegrityError(
            f"Checkpoint at {path} is schema v{man.get('schema_version')}; this release reads "
            f"v{SCHEMA_VERSION}")
    if man.get("sha256") != file_sha256(path):
        raise CheckpointIntegrityError(f"SHA-256 mismatch for {path}")
    if man.get("size_bytes") != path.stat().st_size:
        raise CheckpointIntegrityError(f"size mismatch for {path}")


def load_arrays(path) -> tuple[dict, dict]:
    """``({node: {field: ndarray}}, meta)`` from a checkpoint file."""
    path = Path(path)
    if not path.exists() and path.suffix != ".npz":
        path = path.with_suffix(path.suffix + ".npz")
    if not path.exists():
        raise FileNotFoundError(f"Checkpoint file not found: {path}")
    nodes: dict = {}
    meta: dict = {}
    with np.load(path, allow_pickle=False) as data:
        for key in data.files:
            parts = key.split("/", 1)
            if len(parts) != 2:
                raise ValueError(f"Unexpected key format in checkpoint: '{key}' "
                                 f"(expected 'node_name/field_name')")
            node, field = parts
            (meta if node == _META_KEY else nodes.setdefault(node, {}))[field] = data[key]
    return nodes, meta


def load_state(path, nodes: Mapping[str, object], *, verify: bool = False) -> dict:
    """Restore the states of ``nodes`` (``{name: LBMNode}``) from ``path``.

    The node and field sets must match the file, as in the reference. ``f``
    comes back as the node's SoA-backed ``(*G, Q)`` view on its device, and
    other fields come back as device tensors.
    """
    path = Path(path)
    if verify:
        verify_manifest(path if path.suffix == ".npz" else path.with_suffix(path.suffix + ".npz"))
    saved, _ = load_arrays(path)
    current, stored = set(nodes), set(saved)
    if current != stored:
        parts = []
        if current - stored:
            parts.append(f"missing from checkpoint: {sorted(current - stored)}")
        if stored - current:
            parts.append(f"extra in checkpoint: {sorted(stored - current)}")
        raise ValueError(f"Checkpoint node mismatch. {'; '.join(parts)}")
    out = {}
    for name, node in nodes.items():
        fields = saved[name]
        expect = set(node.initial_state().keys())
        if set(fields) != expect:
            raise ValueError(f"Field mismatch for node '{name}': current={sorted(expect)}, "
                             f"saved={sorted(fields)}")
        k = node.kernel
        st = {}
        for field, arr in fields.items():
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(k.device)
            if field == "f":
                st[field] = k.aos_view(k.from_aos(t))
            elif field == "wall_mask":
                st[field] = t.to(torch.uint8)
            else:
                st[field] = t.to(k.dtype)
        out[name] = st
    return out
