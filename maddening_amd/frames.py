"""Binary state frames in the reference's streaming format, packed on the GPU.

Same layout and API as ``src/maddening/api/binary_encoder.py``:

* The frame is ``[8 B sim_time float64][float32 values]``.
* Nodes and fields appear in sorted order; the ``fields`` argument
  selects a subset, and unknown names are ignored.
* The schema dict has the same keys and ``decode_frame`` inverts
  ``encode``.

A velocity-only subscriber of a 256^3 D3Q19 node therefore receives 8 +
256^3 * 3 * 4 bytes per frame and never sees the 19 populations.

Packing is device-side when the state lives on the GPU. One gather of the
selected fields, float32-cast and flattened in the reference's C order,
goes into a staging buffer, followed by one copy into pinned host memory.
No per-field transfers take place. Compression ("zstd", "zstd+xor") needs
the ``zstandard`` package; when that is absent, asking for it raises
ImportError at construction, as in the reference.
"""
from __future__ import annotations

import struct
from typing import Optional

import numpy as np
import torch

try:  # optional, as in the reference
    import zstandard as _zstd  # type: ignore
    _HAS_ZSTD = True
except ImportError:
    _zstd = None
    _HAS_ZSTD = False

VALID_COMPRESSIONS = {"none", "zstd", "zstd+xor"}


def _shape_of(val) -> tuple:
    if hasattr(val, "shape"):
        return tuple(int(d) for d in val.shape)
    return ()


class BinaryStateEncoder:
    """Fixed schema from an example state; ``encode`` packs later states."""

    def __init__(self, state: dict, fields: Optional[dict] = None, *,
                 compression: str = "none", zstd_level: int = 3):
        if compression not in VALID_COMPRESSIONS:
            raise ValueError(f"compression must be one of {VALID_COMPRESSIONS!r}, "
                             f"got {compression!r}")
        if compression != "none" and not _HAS_ZSTD:
            raise ImportError(f"compression={compression!r} requires the 'zstandard' package")
        self._compression = compression
        self._cctx = _zstd.ZstdCompressor(level=zstd_level) if compression != "none" else None
        self._prev: Optional[bytes] = None
        self._subscription = fields
        self._fields: list[tuple[str, str, tuple, int]] = []
        for node in sorted(state):
            if fields is not None and node not in fields:
                continue
            allowed = None if fields is None else set(fields[node])
            for field in sorted(state[node]):
                if allowed is not None and field not in allowed:
                    continue
                shape = _shape_of(state[node][field])
                n = max(1, int(np.prod(shape))) if shape else 1
                self._fields.append((node, field, shape, n))
        self._total = sum(f[3] for f in self._fields)
        self._staging: dict = {}   # device -> (device buffer, pinned host buffer)

    # -- schema ----------------------------------------------------------------
    def schema(self) -> dict:
        out, off = [], 0
        for node, field, shape, n in self._fields:
            out.append({"node": node, "field": field, "shape": list(shape), "offset": off,
                        "count": n})
            off += n
        return {"type": "schema", "fields": out, "total_floats": self._total,
                "frame_bytes": self.frame_bytes, "compression": self._compression}

    @property
    def total_floats(self) -> int:
        return self._total

    @property
    def frame_bytes(self) -> int:
        return 8 + 4 * self._total

    @property
    def subscription(self):
        return self._subscription

    @property
    def compression(self) -> str:
        return self._compression

    # -- packing -----------------------------------------------------------------
    def _payload(self, state: dict) -> bytes:
        if not self._fields:
            return b""
        vals = [state[node][field] for node, field, _, _ in self._fields]
        dev = next((v.device for v in vals if torch.is_tensor(v) and v.is_cuda), None)
        if dev is None:
            flat = [np.asarray(v.detach().cpu() if torch.is_tensor(v) else v,
                               dtype=np.float32).reshape(-1) for v in vals]
            return np.concatenate(flat).astype(np.float32).tobytes() if flat else b""
        if dev not in self._staging:
            self._staging[dev] = (torch.empty(self._total, dtype=torch.float32, device=dev),
                                  torch.empty(self._total, dtype=torch.float32, pin_memory=True))
        stage, host = self._staging[dev]
        parts = []
        for v in vals:
            t = torch.as_tensor(v, device=dev)
            parts.append(t.reshape(-1).to(torch.float32) if t.dim() else t.reshape(1).float())
        torch.cat(parts, out=stage)
        host.copy_(stage, non_blocking=True)
        torch.cuda.current_stream(dev).synchronize()
        return host.numpy().tobytes()

    def encode(self, sim_time: float, state: dict) -> bytes:
        """``[8 B sim_time][payload]``; the payload is zstd-compressed (after
        an XOR against the previous payload for "zstd+xor") when asked."""
        payload = self._payload(state)
        if self._compression == "zstd":
            payload = self._cctx.compress(payload)
        elif self._compression == "zstd+xor":
            raw = payload
            if self._prev is not None and len(self._prev) == len(raw):
                raw_x = (np.frombuffer(raw, np.uint8) ^ np.frombuffer(self._prev, np.uint8))
                payload = self._cctx.compress(raw_x.tobytes())
            else:
                payload = self._cctx.compress(raw)
            self._prev = raw
        return struct.pack("d", float(sim_time)) + payload


def decode_frame(frame: bytes, schema: dict) -> tuple[float, np.ndarray]:
    """``(sim_time, flat float32 values)`` of a frame ("none" or "zstd")."""
    sim_time = struct.unpack_from("d", frame, 0)[0]
    body = frame[8:]
    comp = schema.get("compression", "none")
    if comp == "zstd":
        if not _HAS_ZSTD:
            raise ImportError("zstandard not installed; cannot decode zstd-compressed frame")
        body = _zstd.ZstdDecompressor().decompress(body)
    elif comp == "zstd+xor":
        raise NotImplementedError("zstd+xor decoding is stateful; the caller keeps the "
                                  "previous payload and XORs after decompressing")
    elif comp != "none":
        raise ValueError(f"unknown compression {comp!r}")
    return sim_time, np.frombuffer(body, dtype=np.float32)


def field_view(values: np.ndarray, schema: dict, node: str, field: str) -> np.ndarray:
    """The slice of decoded ``values`` holding ``node.field``, in its shape."""
    for f in schema["fields"]:
        if f["node"] == node and f["field"] == field:
            seg = values[f["offset"]: f["offset"] + f["count"]]
            return seg.reshape(f["shape"]) if f["shape"] else seg
    raise KeyError(f"{node}.{field} is not in the schema")
