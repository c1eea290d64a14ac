"""Committed golden fixtures (tests/golden/lbm_golden.npz, written by
tests/golden/make_golden.py from the FP64 restatement and cross-checked
against the C restatement): seeded D2Q9 16x12 and D3Q19 8x6x5 states after
1, 10 and 100 steps of lbm.py:899-986. The case set covers periodic
streaming, walls, uniform and per-cell force, a per-cell tau field, the
halfway-wall cavity with a moving lid, and Zou-He faces.

* CPU: both oracles reproduce the fixtures, which pins the oracles against drift.
* GPU: the HIP kernels reproduce them through the C ABI, the f64 kernel to
  1e-12 relative and the f32 kernel to 1e-5.
"""
import os

import numpy as np
import pytest

from oracle import lbm_c as oc
from oracle import lbm_numpy as on

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lbm_golden.npz")
STEPS = (1, 10, 100)


def load_cases():
    with np.load(PATH, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    cases = {}
    for key, val in data.items():
        name, field = key.split("/", 1)
        cases.setdefault(name, {})[field] = val
    return cases


CASES = load_cases()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def faces(c):
    out = []
    for k in ("inlet", "outlet"):
        out.append((str(c[f"{k}_face"]), float(c[f"{k}_pressure"])) if f"{k}_face" in c else None)
    return out


def test_fixture_set_is_complete():
    assert set(CASES) == {"d2q9_periodic_force", "d2q9_walls_force_field",
                          "d2q9_cavity_halfway_lid", "d2q9_zou_he_x_faces",
                          "d3q19_periodic_tau_field", "d3q19_walls_force"}
    for c in CASES.values():
        for n in STEPS:
            assert c[f"f_{n}"].shape == c["f0"].shape
            assert np.isfinite(c[f"f_{n}"]).all()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracles_reproduce_fixtures(name):
    c = CASES[name]
    tau = c["tau_field"] if "tau_field" in c else float(c["tau"])
    inlet, outlet = faces(c)
    kw = dict(mask=c.get("mask"), force=c.get("force"),
              bb="halfway" if "lid" in c else "reference",
              lid_velocity=tuple(c["lid"]) + (0.0,) if "lid" in c else (0.0, 0.0, 0.0),
              inlet=inlet, outlet=outlet)
    res = on.run(c["f0"], 100, tau=tau, **kw)
    assert rel(res["f"], c["f_100"]) < 1e-12
    cres = oc.run(c["f0"], 10, tau=float(c["tau"]), tau_field=c.get("tau_field"), **kw)
    assert rel(cres["f"], c["f_10"]) < 1e-12
    assert rel(cres["density"], c["density_10"]) < 1e-12


def hip_run(c, storage):
    import torch
    from maddening_amd.lbm import LBMKernel
    f0 = c["f0"]
    shape = f0.shape[:-1]
    lat = "D3Q19" if f0.shape[-1] == 19 else "D2Q9"
    bb = "halfway" if "lid" in c else ("reference" if "mask" in c else "none")
    lid = tuple(c["lid"]) + (0.0,) * (3 - len(c["lid"])) if "lid" in c else (0.0, 0.0, 0.0)
    k = LBMKernel(shape, lattice=lat, tau=float(c["tau"]), storage=storage, bb=bb,
                  lid_velocity=lid)
    dev = k.device
    mask = torch.from_numpy(c["mask"]).to(dev) if "mask" in c else None
    force = None
    if "force" in c:
        fr = c["force"]
        force = [float(v) for v in fr] if fr.ndim == 1 else torch.from_numpy(fr).float().to(dev)
    tf = torch.from_numpy(c["tau_field"]).float().to(dev) if "tau_field" in c else None
    fc = [None if s is None else (s[0], s[1] * 3.0) for s in faces(c)]
    fc = None if fc == [None, None] else fc
    f = k.from_aos(torch.from_numpy(f0).to(dev, k.dtype))
    done, out = 0, {}
    for n in STEPS:
        f, rho, u, _ = k.step(f, n - done, mask=mask, force=force, tau_field=tf, faces=fc)
        done = n
        out[n] = (k.to_aos(f, torch.float64).cpu().numpy(), rho.double().cpu().numpy(),
                  u.double().cpu().numpy())
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_f64_reproduces_fixtures(name):
    c = CASES[name]
    out = hip_run(c, "f64")
    for n in STEPS:
        f, rho, u = out[n]
        assert rel(f, c[f"f_{n}"]) < 1e-12, (n, rel(f, c[f"f_{n}"]))
        # moments leave the ABI as float32
        assert rel(rho, c[f"density_{n}"]) < 1e-6
        assert np.max(np.abs(u - c[f"velocity_{n}"])) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_f32_reproduces_fixtures(name):
    c = CASES[name]
    out = hip_run(c, "f32")
    for n in STEPS:
        f, rho, u = out[n]
        assert rel(f, c[f"f_{n}"]) < 1e-5, (n, rel(f, c[f"f_{n}"]))
        assert rel(rho, c[f"density_{n}"]) < 1e-5
        assert np.max(np.abs(u - c[f"velocity_{n}"])) < 2e-6
