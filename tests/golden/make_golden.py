"""Write tests/golden/lbm_golden.npz: seeded inputs and FP64 outputs of the
reference LBM step (lbm.py:899-986) after 1, 10 and 100 steps.

The reference ships no fixtures for this path and cannot run here (JAX is
absent), so the vectors come from the FP64 NumPy restatement
(oracle/lbm_numpy.py). Each output is cross-checked against the independent
C restatement (oracle/lbm_ref.c) before it is written. Cases: D2Q9 16x12 and
D3Q19 8x6x5, with and without walls and force, plus the halfway-wall cavity
and Zou-He pressure faces.

Run from the repository root: ``python tests/golden/make_golden.py``.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import lbm_c as oc  # noqa: E402
from oracle import lbm_numpy as on  # noqa: E402

STEPS = (1, 10, 100)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbm_golden.npz")


def cases():
    rng = np.random.default_rng(2024)

    def f_near_eq(shape, Q):
        e, w, _ = on.lattice(Q)
        rho = 1 + 0.01 * rng.standard_normal(shape)
        u = 0.03 * rng.standard_normal(shape + (len(shape),))
        return on.equilibrium(rho, u, e, w) + 1e-3 * rng.standard_normal(shape + (Q,))

    def f32_exact(x):
        # per-cell fields cross the C ABI as float32: keep them representable
        return np.asarray(x, np.float32).astype(np.float64)

    s2, s3 = (16, 12), (8, 6, 5)
    out = {}
    out["d2q9_periodic_force"] = dict(f0=f_near_eq(s2, 9), tau=0.8, force=np.array([2e-4, -1e-4]))
    out["d2q9_walls_force_field"] = dict(
        f0=f_near_eq(s2, 9), tau=0.7, mask=(rng.random(s2) < 0.2).astype(np.uint8),
        force=f32_exact(1e-4 * rng.standard_normal(s2 + (2,))))
    cav = np.zeros(s2, np.uint8)
    cav[0, :] = cav[-1, :] = cav[:, 0] = 1
    cav[1:-1, -1] = 2
    e9, w9, _ = on.lattice(9)
    out["d2q9_cavity_halfway_lid"] = dict(
        f0=on.equilibrium(np.ones(s2), np.zeros(s2 + (2,)), e9, w9), tau=0.6, mask=cav,
        bb="halfway", lid=np.array([0.1, 0.0]))
    ch = np.zeros(s2, np.uint8)
    ch[:, 0] = ch[:, -1] = 1
    out["d2q9_zou_he_x_faces"] = dict(
        f0=on.equilibrium(np.ones(s2), np.zeros(s2 + (2,)), e9, w9), tau=0.8, mask=ch,
        inlet=("x_min", 1.0 / 3.0 * 1.01), outlet=("x_max", 1.0 / 3.0))
    out["d3q19_periodic_tau_field"] = dict(f0=f_near_eq(s3, 19), tau=0.8,
                                           tau_field=f32_exact(rng.uniform(0.6, 1.3, s3)))
    out["d3q19_walls_force"] = dict(
        f0=f_near_eq(s3, 19), tau=0.9, mask=(rng.random(s3) < 0.2).astype(np.uint8),
        force=np.array([1e-4, 0.0, -2e-4]))
    return out


def kwargs(c):
    kw = {"tau": c.get("tau_field", c["tau"]), "mask": c.get("mask"), "bb": c.get("bb", "reference")}
    if "force" in c:
        kw["force"] = c["force"]
    if "lid" in c:
        kw["lid_velocity"] = tuple(c["lid"]) + (0.0,)
    for k in ("inlet", "outlet"):
        if k in c:
            kw[k] = c[k]
    return kw


def main():
    arrays = {}
    for name, c in cases().items():
        arrays[f"{name}/f0"] = c["f0"]
        arrays[f"{name}/tau"] = np.float64(c["tau"])
        for key in ("mask", "force", "tau_field", "lid"):
            if key in c:
                arrays[f"{name}/{key}"] = c[key]
        for key in ("inlet", "outlet"):
            if key in c:
                arrays[f"{name}/{key}_face"] = np.array(c[key][0])
                arrays[f"{name}/{key}_pressure"] = np.float64(c[key][1])
        kw = kwargs(c)
        f = c["f0"]
        done = 0
        for n in STEPS:
            res = on.run(f, n - done, **kw)
            f, done = res["f"], n
            ckw = dict(kw)
            tf = ckw.pop("tau")
            if np.ndim(tf):
                ckw["tau_field"], ckw["tau"] = tf, c["tau"]
            else:
                ckw["tau"] = tf
            chk = oc.run(c["f0"], n, **ckw)
            err = np.max(np.abs(chk["f"] - res["f"]))
            assert err < 1e-12, (name, n, err)
            arrays[f"{name}/f_{n}"] = res["f"]
            arrays[f"{name}/density_{n}"] = res["density"]
            arrays[f"{name}/velocity_{n}"] = res["velocity"]
    np.savez_compressed(OUT, **arrays)
    print(OUT, os.path.getsize(OUT), "bytes,", len(arrays), "arrays")


if __name__ == "__main__":
    main()
