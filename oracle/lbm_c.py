"""ctypes wrapper of the C restatement (oracle/lbm_ref.c): the second oracle.

Test infrastructure only (tests/ and bench.py's CPU baseline). Arrays are
NumPy, in the reference's layout (*grid_shape, Q), float64 or float32.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .build import LIB, build

FACES = {None: 0, "x_min": 1, "x_max": 2, "y_min": 3, "y_max": 4, "z_min": 5, "z_max": 6}
CS2 = 1.0 / 3.0
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB):
            build()   # one gcc call; the checker may build itself
        h = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        for sfx in ("f64", "f32"):
            common = [ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, ctypes.c_int,
                      ctypes.c_double, vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                      ctypes.c_double, vp, vp, vp, ctypes.c_int]
            getattr(h, f"lbm_oracle_update_{sfx}").argtypes = common
            getattr(h, f"lbm_oracle_run_{sfx}").argtypes = common + [ctypes.c_int]
            getattr(h, f"lbm_oracle_update_{sfx}").restype = ctypes.c_int
            getattr(h, f"lbm_oracle_run_{sfx}").restype = ctypes.c_int
        assert h.lbm_oracle_version() == 1
        _LIB = h
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def run(f, nsteps, *, tau, force=None, mask=None, bb="reference", lid_velocity=(0.0, 0.0, 0.0),
        tau_field=None, inlet=None, outlet=None, nthreads=0, buffers=None):
    """nsteps updates of an AoS state; returns dict f, density, velocity,
    pressure like oracle.lbm_numpy.run. ``force``: None, (D,) or (*G, D);
    ``inlet``/``outlet``: (face name, pressure) or None. ``buffers``: optional
    three arrays shaped like f (state, partner, scratch) reused across calls;
    the first is overwritten with f."""
    f = np.ascontiguousarray(f)
    dt = f.dtype
    if dt not in (np.float64, np.float32):
        raise TypeError("float64 or float32 states only")
    sfx = "f64" if dt == np.float64 else "f32"
    G = f.shape[:-1]
    Q = f.shape[-1]
    D = len(G)
    dims = (ctypes.c_longlong * 3)(*(list(G) + [1] * (3 - D)))
    if buffers is None:
        a, b, post = f.copy(), np.empty_like(f), np.empty_like(f)
    else:
        a, b, post = buffers
        if any(x.shape != f.shape or x.dtype != f.dtype or not x.flags.c_contiguous
               for x in (a, b, post)):
            raise ValueError("buffers must be three C-contiguous arrays shaped like f")
        if a is not f:
            a[...] = f
    ff, field = None, 0
    if force is not None:
        force = np.asarray(force, dtype=dt)
        field = int(force.ndim > 1)
        ff = np.ascontiguousarray(force)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    tf = None if tau_field is None else np.ascontiguousarray(tau_field, dtype=dt)
    lid = (ctypes.c_double * 3)(*(list(lid_velocity) + [0.0] * 3)[:3])
    rho, u, p = np.empty(G, dt), np.empty(G + (D,), dt), np.empty(G, dt)
    fin, fout = (FACES[None], 0.0) if inlet is None else (FACES[inlet[0]], inlet[1] / CS2)
    oin, oout = (FACES[None], 0.0) if outlet is None else (FACES[outlet[0]], outlet[1] / CS2)
    rc = getattr(lib(), f"lbm_oracle_run_{sfx}")(
        Q, dims, _p(a), _p(b), _p(post), _p(m), {"reference": 1, "halfway": 2}[bb], lid, _p(ff),
        field, float(tau), _p(tf), fin, fout, oin, oout, _p(rho), _p(u), _p(p), int(nthreads),
        int(nsteps))
    if rc != 0:
        raise ValueError(f"lbm_oracle_run_{sfx} rejected its arguments ({rc})")
    return {"f": b if nsteps % 2 else a, "density": rho, "velocity": u, "pressure": p}
