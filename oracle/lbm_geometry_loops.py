"""Scalar restatement of the reference's vessel voxelizer: ORACLE.

Test infrastructure only: nothing in the product imports it; only tests/
use it, as the checker. One cell at a time with Python floats (IEEE
float64) and math.sqrt, following src/maddening/nodes/lbm_geometry.py:17-142
step by step: the parent's projection t = clip(x / max(L, 1e-10), 0, 1)
(:86), each daughter's t = clip((P - B) . d / max(|d|^2, 1e-10), 0, 1)
(:106-110, :126-130), closest point B + t d, Euclidean distance, fluid when
strictly below the radius (:93, :117, :137), walls are the rest (:139-143).

Pinned (tests/test_geometry.py) by the reference's own tests
(tests/nodes/test_lbm.py:460-492: fluid fraction 5-50 %, parent centre
fluid) and by closed forms on grids whose section centre is a lattice
point, where the distances are integers: a cell at exactly the radius is a
wall, one cell closer is fluid, along the parent, along straight daughters
and past their rounded ends.
"""
from __future__ import annotations

import math

import numpy as np


def _sq(v):
    """v**2 as numpy evaluates it for float64 arrays: one correctly rounded product."""
    return v * v


def _clip01(v):
    return 0.0 if v < 0.0 else (1.0 if v > 1.0 else v)


def vessel_mask(grid_shape, p):
    nx, ny, nz = grid_shape
    pr, dr = float(p["parent_radius"]), float(p["daughter_radius"])
    pl, dl = float(p["parent_length"]), float(p["daughter_length"])
    a = np.deg2rad(np.float64(p["bifurcation_angle"]))
    ym, zm = (ny - 1) / 2.0, (nz - 1) / 2.0
    segs = []
    # numpy's cos and sin, as the reference evaluates them (lbm_geometry.py:71-72)
    c, s_ = float(np.cos(np.float64(a))), float(np.sin(np.float64(a)))
    for d in ((c, -s_, 0.0), (c, s_, 0.0)):
        segs.append((dl * d[0], dl * d[1], dl * d[2]))
    wall = np.ones((nx, ny, nz), dtype=bool)
    for i in range(nx):
        x = float(i)
        t = _clip01(x / max(pl, 1e-10))
        dxp = x - t * pl
        for j in range(ny):
            y = float(j)
            for k in range(nz):
                z = float(k)
                if math.sqrt(_sq(dxp) + _sq(y - ym) + _sq(z - zm)) < pr:
                    wall[i, j, k] = False
                    continue
                for s in segs:
                    ss = s[0] * s[0] + s[1] * s[1] + s[2] * s[2]
                    t = _clip01(((x - pl) * s[0] + (y - ym) * s[1] + (z - zm) * s[2])
                                / max(ss, 1e-10))
                    d = math.sqrt(_sq(x - (pl + t * s[0])) + _sq(y - (ym + t * s[1]))
                                  + _sq(z - (zm + t * s[2])))
                    if d < dr:
                        wall[i, j, k] = False
                        break
    return wall
