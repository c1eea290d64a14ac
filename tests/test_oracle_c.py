"""The C restatement (oracle/lbm_ref.c) against the NumPy restatement
(oracle/lbm_numpy.py): two independent transcriptions of lbm.py:899-986 must
agree to rounding on every feature the GPU path covers: periodic streaming,
uniform and per-cell force, per-cell tau, the reference wall rule, halfway
bounce-back with a moving lid, and Zou-He faces. The library needs no GPU."""
import numpy as np
import pytest

from oracle import lbm_c as oc
from oracle import lbm_numpy as on


def near_eq(shape, Q, rng, amp=1e-3):
    e, w, _ = on.lattice(Q)
    D = len(shape)
    rho = 1 + 0.01 * rng.standard_normal(shape)
    u = 0.02 * rng.standard_normal(shape + (D,))
    return on.equilibrium(rho, u, e, w) + amp * rng.standard_normal(shape + (Q,))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = {
    "d3q19_periodic_uniform_force": dict(shape=(7, 6, 5), Q=19, tau=0.8, force=[1e-4, -2e-4, 5e-5]),
    "d3q19_walls_force_tau_fields": dict(shape=(9, 10, 11), Q=19, tau=0.8, walls=0.2,
                                         force_field=True, tau_field=True),
    "d2q9_walls_force_field": dict(shape=(13, 11), Q=9, tau=0.7, walls=0.25, force_field=True),
    "d2q9_cavity_halfway": dict(shape=(16, 12), Q=9, tau=0.6, cavity=True),
    "d3q19_halfway_lid": dict(shape=(8, 7, 6), Q=19, tau=0.9, halfway_random=True),
    "d2q9_zou_he_channel": dict(shape=(20, 9), Q=9, tau=0.8, channel=True,
                                inlet=("x_min", 1.0 / 3.0 * 1.01), outlet=("x_max", 1.0 / 3.0)),
    "d3q19_zou_he_two_axes": dict(shape=(6, 7, 8), Q=19, tau=0.75, walls=0.1,
                                  inlet=("z_min", 0.34), outlet=("y_max", 0.33)),
}


def build_case(spec, seed=0):
    rng = np.random.default_rng(seed)
    shape, Q = spec["shape"], spec["Q"]
    D = len(shape)
    kw = {"tau": spec["tau"]}
    f0 = near_eq(shape, Q, rng)
    if "force" in spec:
        kw["force"] = np.array(spec["force"])
    if spec.get("force_field"):
        kw["force"] = 1e-4 * rng.standard_normal(shape + (D,))
    if spec.get("tau_field"):
        kw["tau_field"] = rng.uniform(0.6, 1.3, shape)
    mask = None
    if spec.get("walls"):
        mask = (rng.random(shape) < spec["walls"]).astype(np.uint8)
    if spec.get("cavity"):
        from maddening_amd import workloads
        mask = workloads.cavity_mask(*shape)
        kw["bb"] = "halfway"
        kw["lid_velocity"] = (0.0, 0.1, 0.0)
    if spec.get("halfway_random"):
        mask = (rng.random(shape) < 0.2).astype(np.uint8)
        mask[mask.nonzero()] = rng.integers(1, 3, int(mask.sum())).astype(np.uint8)
        kw["bb"] = "halfway"
        kw["lid_velocity"] = (0.03, -0.02, 0.01)
    if spec.get("channel"):
        mask = np.zeros(shape, np.uint8)
        mask[:, 0] = mask[:, -1] = 1
    for key in ("inlet", "outlet"):
        if key in spec:
            kw[key] = spec[key]
    return f0, mask, kw


def numpy_run(f0, mask, kw, steps):
    kw = dict(kw)
    tf = kw.pop("tau_field", None)
    if tf is not None:
        kw["tau"] = tf
    return on.run(f0, steps, mask=mask, **kw)


@pytest.mark.parametrize("name", sorted(CASES))
def test_c_matches_numpy_float64(name):
    f0, mask, kw = build_case(CASES[name])
    a = numpy_run(f0, mask, kw, 6)
    b = oc.run(f0, 6, mask=mask, nthreads=2, **kw)
    for key in ("f", "density", "velocity", "pressure"):
        assert rel(b[key], a[key]) < 1e-13, (key, rel(b[key], a[key]))


@pytest.mark.parametrize("name", ["d3q19_walls_force_tau_fields", "d2q9_cavity_halfway",
                                  "d2q9_zou_he_channel"])
def test_c_matches_numpy_float32(name):
    """Both restatements evaluated in float32 (the reference's precision)
    stay within a few float32 ulps of each other over 6 steps."""
    f0, mask, kw = build_case(CASES[name])
    kw32 = {k: (np.asarray(v, np.float32) if isinstance(v, np.ndarray) else v)
            for k, v in kw.items()}
    a = numpy_run(f0.astype(np.float32), mask, kw32, 6)
    b = oc.run(f0.astype(np.float32), 6, mask=mask, nthreads=2, **kw32)
    assert b["f"].dtype == np.float32
    assert rel(b["f"], a["f"]) < 2e-6
    assert np.max(np.abs(b["velocity"] - a["velocity"])) < 2e-6


def test_thread_count_does_not_change_results():
    f0, mask, kw = build_case(CASES["d3q19_walls_force_tau_fields"])
    a = oc.run(f0, 4, mask=mask, nthreads=1, **kw)
    b = oc.run(f0, 4, mask=mask, nthreads=4, **kw)
    for key in a:
        assert np.array_equal(a[key], b[key])


def test_rejects_bad_arguments():
    f0, _, _ = build_case(CASES["d2q9_walls_force_field"])
    with pytest.raises(ValueError):
        oc.run(f0, 1, tau=0.8, outlet=("z_max", 0.3))   # no z axis in 2-D
