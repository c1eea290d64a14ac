"""Synthetic inputs of the benchmark and test workloads.

All generators are deterministic (seeded where random) and produce host
NumPy arrays in the reference's layouts: scalar fields (*G), vector fields
(*G, D), masks uint8 (*G) with 1 = wall and 2 = moving wall.
"""
from __future__ import annotations

import numpy as np


def taylor_green_3d(n: int, u0: float, *, with_pressure: bool = True):
    """Taylor-Green vortex on a periodic n^3 grid, k = 2 pi / n.

    u = u0 sin(kx) cos(ky) cos(kz), v = -u0 cos(kx) sin(ky) cos(kz), w = 0,
    rho = 1 + (u0^2/16)(cos 2kx + cos 2ky)(cos 2kz + 2) / cs2 (or 1).
    Axes (0, 1, 2) are (x, y, z). Returns (rho (n,n,n), u (n,n,n,3)) float64.
    """
    k = 2.0 * np.pi / n
    x = np.arange(n, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.empty((n, n, n, 3))
    u[..., 0] = u0 * np.sin(k * X) * np.cos(k * Y) * np.cos(k * Z)
    u[..., 1] = -u0 * np.cos(k * X) * np.sin(k * Y) * np.cos(k * Z)
    u[..., 2] = 0.0
    if with_pressure:
        dp = (u0 ** 2 / 16.0) * (np.cos(2 * k * X) + np.cos(2 * k * Y)) * (np.cos(2 * k * Z) + 2.0)
        rho = 1.0 + 3.0 * dp
    else:
        rho = np.ones((n, n, n))
    return rho, u


def taylor_green_2d(nx: int, ny: int, u0: float):
    """2-D Taylor-Green vortex: u = -u0 cos(kx x) sin(ky y) sqrt(ky/kx),
    v = u0 sin(kx x) cos(ky y) sqrt(kx/ky), rho = 1 - 3 u0^2/4 (...)."""
    kx, ky = 2.0 * np.pi / nx, 2.0 * np.pi / ny
    X, Y = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64),
                       indexing="ij")
    u = np.empty((nx, ny, 2))
    u[..., 0] = -u0 * np.sqrt(ky / kx) * np.cos(kx * X) * np.sin(ky * Y)
    u[..., 1] = u0 * np.sqrt(kx / ky) * np.sin(kx * X) * np.cos(ky * Y)
    p = -0.25 * u0 ** 2 * ((ky / kx) * np.cos(2 * kx * X) + (kx / ky) * np.cos(2 * ky * Y))
    return 1.0 + 3.0 * p, u


def cavity_mask(n0: int, n1: int) -> np.ndarray:
    """Lid-driven cavity: solid walls on axis-0 faces and the axis-1 = 0 face,
    moving lid (value 2) on axis-1 = n1-1 (the corners stay static walls)."""
    m = np.zeros((n0, n1), dtype=np.uint8)
    m[:, n1 - 1] = 2
    m[0, :] = 1
    m[n0 - 1, :] = 1
    m[:, 0] = 1
    return m


def channel_walls(shape) -> np.ndarray:
    """Walls on the axis-0 and axis-1 faces of a 3-D box (axis 2 periodic)."""
    m = np.zeros(shape, dtype=np.uint8)
    m[0, :, :] = m[-1, :, :] = 1
    m[:, 0, :] = m[:, -1, :] = 1
    return m


def sphere_packing(shape, *, porosity: float = 0.7, rmin: int = 6, rmax: int = 12,
                   seed: int = 0, walls: bool = True) -> np.ndarray:
    """Seeded random sphere packing, integer-exact.

    Spheres with integer centres drawn uniformly in the box and integer radii
    uniform in [rmin, rmax] are added (overlaps allowed, periodic along the
    last axis) until the fluid fraction drops to ``porosity`` or below.
    Inside test (i - c)^2 + ... <= r^2 in int64, so the mask is bit-exact on
    every platform. Optional channel walls on the first two axes.
    """
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    m = channel_walls(shape) if walls else np.zeros(shape, dtype=np.uint8)
    total = int(np.prod(shape))
    idx = [np.arange(s, dtype=np.int64) for s in shape]
    while np.count_nonzero(m == 0) / total > porosity:
        c = [int(rng.integers(0, s)) for s in shape]
        r = int(rng.integers(rmin, rmax + 1))
        lo = [max(ci - r, 0) for ci in c[:2]]
        hi = [min(ci + r + 1, s) for ci, s in zip(c[:2], shape[:2])]
        d0 = (idx[0][lo[0]:hi[0]] - c[0]) ** 2
        d1 = (idx[1][lo[1]:hi[1]] - c[1]) ** 2
        # periodic distance along the last axis
        dz = np.abs(idx[2] - c[2])
        dz = np.minimum(dz, shape[2] - dz) ** 2
        inside = (d0[:, None, None] + d1[None, :, None] + dz[None, None, :]) <= r * r
        sub = m[lo[0]:hi[0], lo[1]:hi[1], :]
        sub[inside & (sub == 0)] = 1
    return m


def batch_specs(seed: int = 0, n2d: int = 32, n3d: int = 32):
    """Uneven batch: n2d D2Q9 squares with edge U[64,1024], n3d D3Q19 cubes
    with edge U[32,160], tau U[0.55,1.5], Taylor-Green or cavity init."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n2d):
        n = int(rng.integers(64, 1025))
        specs.append({"lattice": "D2Q9", "shape": (n, n), "tau": float(rng.uniform(0.55, 1.5)),
                      "init": "cavity" if rng.random() < 0.5 else "taylor_green"})
    for _ in range(n3d):
        n = int(rng.integers(32, 161))
        specs.append({"lattice": "D3Q19", "shape": (n, n, n),
                      "tau": float(rng.uniform(0.55, 1.5)), "init": "taylor_green"})
    return specs
