"""Wall masks for the LBM node (host side, float64, one-off set-up).

``voxelize_vessel`` restates ``src/maddening/nodes/lbm_geometry.py:17-142``:
the Y-bifurcation vessel of the vessel-flow examples (a parent tube along x
that splits into two daughter tubes in the x-y plane). A cell is fluid when
its distance to any of the three centreline segments is strictly below that
segment's radius; the mask is True at walls. The distances are evaluated in
float64 with the reference's operations, in its order (the parent's
projection onto its axis-aligned segment is x / L, the daughters' the dot
product over |d|^2), so cells that sit exactly on a radius land on the same
side of it.
"""
from __future__ import annotations

from typing import Optional

import numpy as np


def _daughter_fluid(xx, yy, zz, start, seg, radius):
    """Cells closer than `radius` to the segment start .. start + seg."""
    dx, dy, dz = xx - start[0], yy - start[1], zz - start[2]
    t = np.clip((dx * seg[0] + dy * seg[1] + dz * seg[2]) / max(float(np.dot(seg, seg)), 1e-10),
                0.0, 1.0)
    cx = start[0] + t * seg[0]
    cy = start[1] + t * seg[1]
    cz = start[2] + t * seg[2]
    return np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) < radius


def voxelize_vessel(grid_shape: tuple, vessel_params: dict, device: Optional[str] = None):
    """Boolean wall mask (True = wall) of a Y-bifurcation vessel
    (lbm_geometry.py:17-142).

    ``vessel_params``: ``parent_radius``, ``daughter_radius``,
    ``parent_length``, ``daughter_length`` (grid units) and
    ``bifurcation_angle`` (degrees; each daughter leaves the bifurcation
    point at +-angle from the x axis, towards +y and -y). The parent runs
    from x = 0 to x = parent_length along the centre of the (y, z) section.

    Returns a numpy bool array of ``grid_shape``, or a torch uint8 tensor on
    ``device`` when one is given (the node's mask layout).
    """
    nx, ny, nz = (int(n) for n in grid_shape)
    pr = float(vessel_params["parent_radius"])
    dr = float(vessel_params["daughter_radius"])
    pl = float(vessel_params["parent_length"])
    dl = float(vessel_params["daughter_length"])
    ang = np.deg2rad(float(vessel_params["bifurcation_angle"]))
    ymid, zmid = (ny - 1) / 2.0, (nz - 1) / 2.0
    xx, yy, zz = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64),
                             np.arange(nz, dtype=np.float64), indexing="ij")
    # parent: axis-aligned segment (0, ymid, zmid) .. (pl, ymid, zmid)
    cx = np.clip(xx / max(pl, 1e-10), 0.0, 1.0) * pl
    fluid = np.sqrt((xx - cx) ** 2 + (yy - ymid) ** 2 + (zz - zmid) ** 2) < pr
    bif = np.array([pl, ymid, zmid])
    for sign in (-1.0, 1.0):   # the -y daughter, then the +y daughter
        seg = dl * np.array([np.cos(ang), sign * np.sin(ang), 0.0])
        fluid |= _daughter_fluid(xx, yy, zz, bif, seg, dr)
    wall = ~fluid
    if device is None:
        return wall
    import torch
    return torch.from_numpy(wall.astype(np.uint8)).to(device)
