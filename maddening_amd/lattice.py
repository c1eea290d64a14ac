"""Lattice descriptors with the reference's API (lbm.py:57-173).

``LatticeDescriptor``, ``d3q19()`` and ``d2q9()`` have the same fields,
direction order, weights and opposite map as MADDENING's, because the state
layout, checkpoints and bounce-back link indexing depend on that order. The
device kernels hold the same tables as compile-time constants
(csrc/lbm_common.hpp); tests check that both agree.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class LatticeDescriptor:
    """Velocity set: name, e (Q, D) int32, w (Q,) float64, opp (Q,) int32,
    cs2, D, Q. NumPy arrays so they stay concrete constants."""
    name: str
    e: np.ndarray
    w: np.ndarray
    opp: np.ndarray
    cs2: float
    D: int
    Q: int


def _opp(e: np.ndarray) -> np.ndarray:
    # index of the reversed velocity; the lattices are symmetric under e -> -e
    lookup = {tuple(v): i for i, v in enumerate(e.tolist())}
    return np.array([lookup[tuple(-v for v in row)] for row in e.tolist()], dtype=np.int32)


def d3q19() -> LatticeDescriptor:
    """D3Q19: rest, 6 face neighbours (w 1/18), 12 edge neighbours (w 1/36),
    ordered +x,-x,+y,-y,+z,-z, then xy, xz, yz edges (lbm.py:90-139)."""
    axes = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    edges = []
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for sb in (1, -1):
            for sa in (1, -1):
                v = [0, 0, 0]
                v[a], v[b] = sa, sb
                edges.append(tuple(v))
    e = np.array([(0, 0, 0)] + axes + edges, dtype=np.int32)
    w = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12, dtype=np.float64)
    return LatticeDescriptor(name="D3Q19", e=e, w=w, opp=_opp(e), cs2=1.0 / 3.0, D=3, Q=19)


def d2q9() -> LatticeDescriptor:
    """D2Q9: rest (w 4/9), 4 axis (w 1/9), 4 diagonals (w 1/36), ordered
    +x,-x,+y,-y, then (+,+),(-,+),(+,-),(-,-) (lbm.py:142-173)."""
    axes = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    diags = [(sx, sy) for sy in (1, -1) for sx in (1, -1)]
    e = np.array([(0, 0)] + axes + diags, dtype=np.int32)
    w = np.array([4.0 / 9.0] + [1.0 / 9.0] * 4 + [1.0 / 36.0] * 4, dtype=np.float64)
    return LatticeDescriptor(name="D2Q9", e=e, w=w, opp=_opp(e), cs2=1.0 / 3.0, D=2, Q=9)
