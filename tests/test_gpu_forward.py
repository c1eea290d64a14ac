"""GPU parity: HIP forward kernels vs the float64 oracle (oracle/lbm_numpy.py).

Tolerances (written per test): float64 storage 1e-12 relative, float32
storage 1e-5 relative on rho and u (the float32 bar against float64),
float16/bfloat16 deviation storage 1e-3 / 1e-2 relative velocity error.
Relative errors are max|a - b| / max|b| over the grid.
"""
import numpy as np
import pytest
import torch

from oracle import lbm_numpy as orc
from maddening_amd import workloads

pytestmark = pytest.mark.gpu

SCHEMES = ["pull", "push"]


def kernel(shape, lattice, tau, **kw):
    from maddening_amd.lbm import LBMKernel
    return LBMKernel(shape, lattice=lattice, tau=tau, device="cuda", **kw)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_both(shape, lattice, tau, nsteps, *, rho0=None, u0=None, storage="f32",
             scheme="pull", bb="none", mask=None, force=None, tau_field=None,
             lid=(0.0, 0.0, 0.0), f0_aos=None, steps_per_call=None, row_vec=0, faces=None):
    """Same initial state through the oracle and the HIP kernel."""
    Q = 19 if lattice == "D3Q19" else 9
    e, w, _ = orc.lattice(Q)
    if f0_aos is None:
        rho = np.ones(shape) if rho0 is None else rho0
        u = np.zeros(tuple(shape) + (len(shape),)) if u0 is None else u0
        f0_aos = orc.equilibrium(rho, u, e, w)
    k = kernel(shape, lattice, tau, storage=storage, scheme=scheme, bb=bb, lid_velocity=lid,
               row_vec=row_vec)
    f = k.from_aos(torch.from_numpy(np.ascontiguousarray(f0_aos)).to("cuda", torch.float64))
    mask_t = None if mask is None else torch.from_numpy(mask).cuda()
    force_arg = force
    if isinstance(force, np.ndarray) and force.ndim > 1:
        # the kernel reads float32 fields: give the oracle the same rounded values
        force = force.astype(np.float32).astype(np.float64)
        force_arg = torch.from_numpy(force.astype(np.float32)).cuda()
    tf = None if tau_field is None else torch.from_numpy(tau_field.astype(np.float32)).cuda()
    per = steps_per_call or nsteps
    done = 0
    while done < nsteps:
        n = min(per, nsteps - done)
        f, rho_g, u_g, p_g = k.step(f, n, mask=mask_t, force=force_arg, tau_field=tf,
                                    faces=faces)
        done += n
    torch.cuda.synchronize()
    fo = f0_aos.astype(np.float64)
    tau_o = tau if tau_field is None else tau_field.astype(np.float32).astype(np.float64)
    zh = {}
    if faces is not None:
        if faces[0] is not None:
            zh["inlet"] = (faces[0][0], faces[0][1] / 3.0)   # oracle takes pressure
        if faces[1] is not None:
            zh["outlet"] = (faces[1][0], faces[1][1] / 3.0)
    out = orc.run(fo, nsteps, tau=tau_o, force=force, mask=mask,
                  bb="halfway" if bb == "halfway" else "reference", lid_velocity=lid, **zh)
    gpu = {"f": k.to_aos(f, torch.float64).cpu().numpy(), "density": rho_g.cpu().numpy(),
           "velocity": u_g.cpu().numpy(), "pressure": p_g.cpu().numpy()}
    return gpu, out


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_d3q19_taylor_green_f32(scheme, storage):
    rho, u = workloads.taylor_green_3d(32, 0.05)
    g, o = run_both((32, 32, 32), "D3Q19", 0.8, 100, rho0=rho, u0=u, scheme=scheme,
                    storage=storage)
    assert rel(g["density"], o["density"]) < 1e-5
    # deviation storage rounds |f - w| instead of w: an order of magnitude tighter
    assert rel(g["velocity"], o["velocity"]) < (1e-5 if storage == "f32" else 1e-6)
    assert rel(g["f"], o["f"]) < 1e-5
    np.testing.assert_allclose(g["pressure"], o["density"] / 3.0, rtol=1e-5)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_d3q19_f64_matches_oracle_tightly(scheme):
    rho, u = workloads.taylor_green_3d(16, 0.05)
    g, o = run_both((16, 16, 16), "D3Q19", 0.8, 50, rho0=rho, u0=u, storage="f64", scheme=scheme)
    assert rel(g["f"], o["f"]) < 1e-12
    # moments are returned as float32
    assert rel(g["velocity"], o["velocity"]) < 1e-6


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("shape", [(12, 10, 7), (5, 9, 70)])
def test_d3q19_reference_walls_force_field_f64(scheme, shape):
    rng = np.random.default_rng(3)
    mask = (rng.random(shape) < 0.2).astype(np.uint8)
    force = 1e-4 * rng.standard_normal(shape + (3,))
    tau_field = rng.uniform(0.6, 1.2, shape)
    f0 = orc.equilibrium(1 + 0.01 * rng.standard_normal(shape),
                         0.02 * rng.standard_normal(shape + (3,)), *orc.lattice(19)[:2])
    g, o = run_both(shape, "D3Q19", 0.8, 7, storage="f64", scheme=scheme, bb="reference",
                    mask=mask, force=force, tau_field=tau_field, f0_aos=f0)
    assert rel(g["f"], o["f"]) < 1e-12
    assert np.all(g["velocity"][mask != 0] == 0.0)
    assert rel(g["velocity"], o["velocity"]) < 1e-6


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_d3q19_uniform_force_f32(scheme, storage):
    shape = (16, 12, 20)
    mask = workloads.channel_walls(shape)
    g, o = run_both(shape, "D3Q19", 0.7, 60, storage=storage, scheme=scheme, bb="reference",
                    mask=mask, force=(0.0, 0.0, 1e-5))
    assert rel(g["density"], o["density"]) < 1e-5
    # |u| only reaches ~6e-4 here, while storing f ~ w in float32 rounds each
    # population by up to ~4e-9 per step; that floor, not the kernel, sets the
    # error, so the bound is absolute: 1.5e-7 (3e-6 of a 0.05 velocity scale).
    err = np.max(np.abs(g["velocity"] - o["velocity"]))
    assert err < 1.5e-7, err


@pytest.mark.parametrize("scheme", SCHEMES)
def test_d2q9_cavity_halfway_f32dev(scheme):
    mask = workloads.cavity_mask(24, 20)
    g, o = run_both((24, 20), "D2Q9", 0.6, 40, storage="f32dev", scheme=scheme, bb="halfway",
                    mask=mask, lid=(0.1, 0.0, 0.0))
    assert rel(g["velocity"], o["velocity"]) < 1e-6


@pytest.mark.parametrize("scheme", SCHEMES)
def test_d2q9_cavity_halfway_f64(scheme):
    mask = workloads.cavity_mask(24, 20)
    g, o = run_both((24, 20), "D2Q9", 0.6, 40, storage="f64", scheme=scheme, bb="halfway",
                    mask=mask, lid=(0.1, 0.0, 0.0))
    fluid = mask == 0
    assert rel(g["f"][fluid], o["f"][fluid]) < 1e-12
    assert rel(g["velocity"], o["velocity"]) < 1e-6


@pytest.mark.parametrize("scheme", SCHEMES)
def test_d2q9_periodic_f32_multi_call(scheme):
    rho, u = workloads.taylor_green_2d(40, 24, 0.04)
    g, o = run_both((40, 24), "D2Q9", 0.9, 30, rho0=rho, u0=u, scheme=scheme, steps_per_call=7)
    assert rel(g["density"], o["density"]) < 1e-5
    assert rel(g["velocity"], o["velocity"]) < 1e-5


@pytest.mark.parametrize("storage,tol", [("f16dev", 2e-3), ("bf16dev", 2e-2)])
def test_d3q19_deviation_storage(storage, tol):
    rho, u = workloads.taylor_green_3d(32, 0.05)
    g, o = run_both((32, 32, 32), "D3Q19", 0.8, 100, rho0=rho, u0=u, storage=storage)
    assert rel(g["velocity"], o["velocity"]) < tol


def test_equilibrium_moments_roundtrip():
    from maddening_amd.lbm import LBMKernel
    rho, u = workloads.taylor_green_3d(16, 0.05)
    k = LBMKernel((16, 16, 16), tau=0.8)
    f = k.equilibrium(torch.from_numpy(rho).float(), torch.from_numpy(u).float())
    r, v, p = k.moments(f)
    np.testing.assert_allclose(r.cpu().numpy(), rho, rtol=1e-6)
    np.testing.assert_allclose(v.cpu().numpy(), u, atol=1e-7)
    e, w, _ = orc.lattice(19)
    np.testing.assert_allclose(k.aos_view(f).cpu().numpy(), orc.equilibrium(rho, u, e, w),
                               rtol=1e-6, atol=1e-8)


def test_conservation_sums_f64():
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(0)
    aos = rng.random((10, 11, 12, 19))
    k = LBMKernel((10, 11, 12), tau=0.8, storage="f64")
    f = k.from_aos(torch.from_numpy(aos).cuda())
    s = k.conservation(f).cpu().numpy()
    e, _, _ = orc.lattice(19)
    ref = np.concatenate([[aos.sum()], np.einsum("xyzq,qd->d", aos, e.astype(np.float64))])
    np.testing.assert_allclose(s, ref, rtol=1e-13)


def test_periodic_mass_conserved_f32():
    """Reference test_lbm.py:336-357: 16^3, 50 steps, rtol 1e-5."""
    from maddening_amd.lbm import LBMKernel
    k = LBMKernel((16, 16, 16), tau=0.8)
    rho, u = workloads.taylor_green_3d(16, 0.05)
    f = k.equilibrium(torch.from_numpy(rho).float(), torch.from_numpy(u).float())
    m0 = float(k.conservation(f)[0])
    f, _, _, _ = k.step(f, 50)
    m1 = float(k.conservation(f)[0])
    assert abs(m1 - m0) / m0 < 1e-5


def test_errors_are_reported():
    from maddening_amd.lbm import LBMKernel
    from maddening_amd._lib import LBMError
    k = LBMKernel((8, 8, 8), tau=0.8, bb="reference")
    f = k.equilibrium()
    with pytest.raises(ValueError, match="wall mask"):
        k.step(f, 1)
    with pytest.raises(LBMError, match="distinct"):
        k.step(f, 1, mask=torch.zeros(8, 8, 8, dtype=torch.uint8), out=f)


# --- row-per-wave kernels (periodic, contiguous axis a multiple of 64) -------
ROW_CASES = [
    ("D3Q19", (6, 5, 256), "f32", 4),     # 4 cells per lane, one wave per row
    ("D3Q19", (4, 3, 512), "f32", 4),     # two waves per row: cross-wave edge loads
    ("D3Q19", (5, 4, 192), "f32", 1),     # 1 cell per lane, three waves per row
    ("D3Q19", (4, 4, 256), "f64", 2),     # 2 cells per lane, two waves per row
    ("D3Q19", (4, 4, 128), "f16dev", 2),
    ("D2Q9", (12, 256), "f32", 4),
    ("D2Q9", (7, 640), "bf16dev", 2),     # five waves per row
]


@pytest.mark.parametrize("lattice,shape,storage,vec", ROW_CASES)
def test_row_kernel_matches_cell_kernel_bitwise(lattice, shape, storage, vec):
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(11)
    D = len(shape)
    rho = torch.from_numpy(1 + 0.01 * rng.standard_normal(shape)).float()
    u = torch.from_numpy(0.03 * rng.standard_normal(shape + (D,))).float()
    outs = []
    for row_vec in (vec, -1):
        k = LBMKernel(shape, lattice=lattice, tau=0.7, storage=storage, scheme="pull",
                      row_vec=row_vec)
        f = k.equilibrium(rho, u)
        f, r, v, p = k.step(f, 9, force=[1e-5] * D)
        torch.cuda.synchronize()
        outs.append((f.cpu(), r.cpu(), v.cpu()))
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("lattice,shape,storage,vec",
                         [c for c in ROW_CASES if c[2] in ("f32", "f64")])
def test_row_kernel_matches_oracle(lattice, shape, storage, vec):
    rng = np.random.default_rng(12)
    D = len(shape)
    rho0 = 1 + 0.01 * rng.standard_normal(shape)
    u0 = 0.03 * rng.standard_normal(shape + (D,))
    g, o = run_both(shape, lattice, 0.7, 9, rho0=rho0, u0=u0, storage=storage,
                    scheme="pull", force=tuple([1e-5] * D), row_vec=vec)
    assert rel(g["f"], o["f"]) < (1e-12 if storage == "f64" else 1e-5)
    assert rel(g["velocity"], o["velocity"]) < 1e-5


# --- Zou-He pressure faces (lbm.py:425-524) -----------------------------------
@pytest.mark.parametrize("scheme", SCHEMES)
def test_zou_he_faces_d2q9_channel_f64(scheme):
    shape = (40, 12)
    mask = np.zeros(shape, np.uint8)
    mask[:, 0] = mask[:, -1] = 1
    g, o = run_both(shape, "D2Q9", 0.8, 30, storage="f64", scheme=scheme, bb="reference",
                    mask=mask, faces=(("x_min", 1.01), ("x_max", 1.0)))
    assert rel(g["f"], o["f"]) < 1e-12
    assert rel(g["velocity"], o["velocity"]) < 1e-6


@pytest.mark.parametrize("scheme", SCHEMES)
def test_zou_he_faces_d3q19_other_axes_f64(scheme):
    """Inlet on y_min and outlet on z_max (two different axes, shared edge),
    random walls: checks the axis mapping and the inlet-then-outlet order."""
    shape = (9, 7, 8)
    rng = np.random.default_rng(9)
    mask = (rng.random(shape) < 0.1).astype(np.uint8)
    f0 = orc.equilibrium(1 + 0.01 * rng.standard_normal(shape),
                         0.02 * rng.standard_normal(shape + (3,)), *orc.lattice(19)[:2])
    g, o = run_both(shape, "D3Q19", 0.9, 12, storage="f64", scheme=scheme, bb="reference",
                    mask=mask, f0_aos=f0, faces=(("y_min", 1.02), ("z_max", 0.99)))
    assert rel(g["f"], o["f"]) < 1e-12


def test_zou_he_faces_f32_periodic_row_path_off():
    """Faces on a periodic lattice whose rows could use the row kernel: the
    faces must still be applied (the row kernel is bypassed)."""
    shape = (6, 5, 64)
    g, o = run_both(shape, "D3Q19", 0.7, 10, storage="f32", scheme="pull", row_vec=1,
                    faces=(("z_min", 1.01), None))
    assert rel(g["density"], o["density"]) < 1e-5


@pytest.mark.parametrize("scheme,pair", [("pull", "off"), ("pull", "auto"), ("push", "auto"),
                                         ("auto", "auto")])
@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("lattice,axis", [("D3Q19", 0), ("D3Q19", 1), ("D3Q19", 2),
                                          ("D2Q9", 0), ("D2Q9", 1)])
def test_zou_he_resting_face_closed_form(lattice, axis, storage, scheme, pair):
    """The reference's face rule (lbm.py:476-479, 503) in closed form: a
    lattice at rest at rho_p with both faces of one axis at rho_p. After one
    step every unknown slot of a fluid face cell is exactly 2 w_q rho_p (the
    textbook Zou-He rule would leave w_q rho_p), a wall cell on the face and
    every bulk cell keep w rho_p. Covers the pull sequence's stream + moments
    launch (pair off), the push path's face kernel and the one-launch
    update (D3Q19 float32 pull or auto)."""
    from maddening_amd.lbm import LBMKernel
    Q = 19 if lattice == "D3Q19" else 9
    e, w, _ = orc.lattice(Q)
    shape = (6, 5, 8) if Q == 19 else (6, 8)
    D = len(shape)
    rho_p = 1.02
    f0 = np.broadcast_to(w * rho_p, shape + (Q,)).copy()
    mask = np.zeros(shape, np.uint8)
    idx = [2] * D
    idx[axis] = 0
    mask[tuple(idx)] = 1
    names = LBMKernel.FACES[2 * axis: 2 * axis + 2]
    k = kernel(shape, lattice, 0.8, storage=storage, scheme=scheme, bb="reference", pair=pair)
    f = k.from_aos(torch.from_numpy(f0).cuda())
    fo, _, _, _ = k.step(f, 1, mask=torch.from_numpy(mask).cuda(),
                         faces=((names[0], rho_p), (names[1], rho_p)))
    got = k.to_aos(fo, torch.float64).cpu().numpy()
    expect = f0.copy()
    for side, x in (("min", 0), ("max", shape[axis] - 1)):
        _, unknown, _ = orc.classify_directions(e, axis, side)
        sl = [slice(None)] * D
        sl[axis] = x
        for q in unknown:
            expect[tuple(sl) + (q,)] = np.where(mask[tuple(sl)] != 0, w[q] * rho_p,
                                                2 * w[q] * rho_p)
    tol = 1e-14 if storage == "f64" else 2e-6
    np.testing.assert_allclose(got, expect, rtol=tol, atol=0)


# degenerate and ragged extents: axes of length 1 and 2 (periodic neighbours
# wrap onto the cell itself or onto the same partner from both sides), rows
# shorter than a wave and not a multiple of 64, prime extents
DEGENERATE = [("D3Q19", (1, 1, 1)), ("D3Q19", (1, 5, 3)), ("D3Q19", (2, 1, 7)),
              ("D3Q19", (3, 2, 1)), ("D3Q19", (7, 11, 13)), ("D3Q19", (4, 3, 257)),
              ("D2Q9", (1, 1)), ("D2Q9", (1, 9)), ("D2Q9", (13, 1)), ("D2Q9", (2, 2)),
              ("D2Q9", (5, 131))]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("lattice,shape", DEGENERATE)
def test_degenerate_extents_f64(lattice, shape, scheme):
    rng = np.random.default_rng(sum(shape))
    D = len(shape)
    rho = 1 + 0.02 * rng.standard_normal(shape)
    u = 0.03 * rng.standard_normal(shape + (D,))
    force = 1e-4 * rng.standard_normal(shape + (D,))
    mask = (rng.random(shape) < 0.2).astype(np.uint8)
    for bb, m in (("none", None), ("reference", mask)):
        gpu, ref = run_both(shape, lattice, 0.77, 6, rho0=rho, u0=u, storage="f64",
                            scheme=scheme, bb=bb, mask=m, force=force)
        for key in ("f", "density", "velocity"):
            assert rel(gpu[key], ref[key]) < 1e-12 if key == "f" else \
                rel(gpu[key], ref[key]) < 1e-6, (bb, key, rel(gpu[key], ref[key]))


def _halfway_walls(shape, seed):
    """Random static (1) and moving (2) wall cells."""
    rng = np.random.default_rng(seed)
    r = rng.random(shape)
    return np.where(r < 0.15, 1, np.where(r < 0.25, 2, 0)).astype(np.uint8)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("shape", [(12, 10, 9), (3, 17, 5)])
def test_d3q19_halfway_moving_walls_f64(scheme, shape):
    """Halfway bounce-back with static and moving wall cells scattered through
    the lattice, a per-cell force and a lid velocity on every axis."""
    rng = np.random.default_rng(11)
    mask = _halfway_walls(shape, 5)
    force = 1e-4 * rng.standard_normal(shape + (3,))
    f0 = orc.equilibrium(1 + 0.01 * rng.standard_normal(shape),
                         0.02 * rng.standard_normal(shape + (3,)), *orc.lattice(19)[:2])
    g, o = run_both(shape, "D3Q19", 0.9, 9, storage="f64", scheme=scheme, bb="halfway",
                    mask=mask, force=force, f0_aos=f0, lid=(0.03, -0.02, 0.05))
    fluid = mask == 0
    assert rel(g["f"][fluid], o["f"][fluid]) < 1e-12
    assert np.all(g["velocity"][~fluid] == 0.0)
    assert rel(g["velocity"], o["velocity"]) < 1e-6


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (16, 14, 40)), ("D2Q9", (30, 70))])
@pytest.mark.parametrize("storage", ["f32dev", "f16dev", "f64"])
def test_halfway_link_words_match_mask_path_bitwise(lattice, shape, storage):
    """lbm_step's pull sequence reads per-cell link words built once per call;
    lbm_pull_phase reads the wall mask bytes directly, or the link words of
    lbm_links_build. The same launches in each form give the same bits."""
    import ctypes
    from maddening_amd import _lib
    from maddening_amd.lbm import _ptr, _stream
    D = len(shape)
    rng = np.random.default_rng(7)
    mask = torch.from_numpy(_halfway_walls(shape, 9)).cuda()
    lid = (0.04, 0.01, -0.03)[:D] + (0.0,) * (3 - D)
    k = kernel(shape, lattice, 0.7, storage=storage, scheme="pull", bb="halfway", lid_velocity=lid)
    Q = k.Q
    e, w, _ = orc.lattice(Q)
    f0 = orc.equilibrium(1 + 0.01 * rng.standard_normal(shape),
                         0.03 * rng.standard_normal(shape + (D,)), e, w)
    f = k.from_aos(torch.from_numpy(f0).cuda())
    force = [2e-5, -1e-5, 3e-5][:D]
    nsteps = 6
    f_step, rho_s, u_s, _ = k.step(f, nsteps, mask=mask, force=force)
    L = _lib.lib()
    st = _stream(k.device)
    d = k.desc(force=force)
    links = torch.empty(L.lbm_links_size(ctypes.byref(d)) // 4, dtype=torch.int32, device="cuda")
    _lib.check(L.lbm_links_build(st, ctypes.byref(d), _ptr(mask), _ptr(links)), "links")
    for lk in (None, links):
        a, b = k.empty(), k.empty()
        rho = torch.empty(k.grid_shape, dtype=torch.float32, device="cuda")
        u = torch.empty(k.grid_shape + (D,), dtype=torch.float32, device="cuda")
        _lib.check(L.lbm_pull_phase(st, ctypes.byref(d), 0, _ptr(f), _ptr(a), _ptr(mask), _ptr(lk),
                                    None, None, None, None, None), "phase 0")
        for _ in range(nsteps - 1):
            _lib.check(L.lbm_pull_phase(st, ctypes.byref(d), 1, _ptr(a), _ptr(b), _ptr(mask),
                                        _ptr(lk), None, None, None, None, None), "phase 1")
            a, b = b, a
        out = k.empty()
        _lib.check(L.lbm_pull_phase(st, ctypes.byref(d), 2, _ptr(a), _ptr(out), _ptr(mask),
                                    _ptr(lk), None, None, _ptr(rho), _ptr(u), None), "phase 2")
        torch.cuda.synchronize()
        assert torch.equal(f_step.view(torch.uint8), out.view(torch.uint8))
        assert torch.equal(rho_s, rho) and torch.equal(u_s, u)


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (9, 11, 161)), ("D3Q19", (5, 6, 320)),
                                           ("D3Q19", (7, 5, 33)), ("D2Q9", (37, 193))])
@pytest.mark.parametrize("bb", ["none", "reference", "halfway"])
@pytest.mark.parametrize("storage", ["f32", "f16dev"])
def test_flat_tiles_match_row_blocks_bitwise(lattice, shape, bb, storage):
    """Rows that are not a multiple of 64 cells (or ragged past 256) run over
    256-cell tiles of the flattened lattice; an explicit block_x keeps row
    blocks. Both launch shapes give the same bits."""
    rng = np.random.default_rng(len(shape) + shape[-1])
    D = len(shape)
    Q = 19 if lattice == "D3Q19" else 9
    e, w, _ = orc.lattice(Q)
    f0 = orc.equilibrium(1 + 0.01 * rng.standard_normal(shape),
                         0.03 * rng.standard_normal(shape + (D,)), e, w)
    mask = None
    if bb != "none":
        mask = torch.from_numpy((rng.random(shape) < 0.15).astype(np.uint8) *
                                (2 if bb == "halfway" else 1)).cuda()
    lid = (0.02, -0.01, 0.03)[:D] + (0.0,) * (3 - D)
    outs = []
    for bx in (0, 256):
        k = kernel(shape, lattice, 0.75, storage=storage, bb=bb, lid_velocity=lid, block_x=bx)
        f = k.from_aos(torch.from_numpy(f0).cuda())
        fo, rho, u, _ = k.step(f, 5, mask=mask, force=[1e-5, -2e-5, 3e-5][:D])
        outs.append((fo.view(torch.uint8), rho, u))
    torch.cuda.synchronize()
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape,bx", [((65, 64, 256), 0), ((25, 17, 256), 0), ((9, 11, 161), 0),
                                      ((9, 11, 161), -1), ((40, 32, 64), 64)])
def test_xcd_run_orders_match_bitwise(shape, bx):
    """The XCD remap only permutes which workgroup takes which tile: identity,
    runs of 3 or 512 (with a remainder past the last whole set) and one
    eighth per XCD give the same bits as the library's own choice, for row
    blocks and flattened tiles, forward and through the adjoint."""
    from maddening_amd.adjoint import rollout
    rng = np.random.default_rng(5)
    D = len(shape)
    e, w, _ = orc.lattice(19)
    f0 = orc.equilibrium(1 + 0.01 * rng.standard_normal(shape),
                         0.03 * rng.standard_normal(shape + (D,)), e, w)
    mask = torch.from_numpy((rng.random(shape) < 0.15).astype(np.uint8)).cuda()
    tau = torch.from_numpy(rng.uniform(0.6, 1.1, shape)).float().cuda()
    outs = []
    for remap, G in ((False, 0), (True, 1), (True, 3), (True, 512), (True, 0)):
        k = kernel(shape, "D3Q19", 0.8, storage="f32", bb="reference", block_x=bx,
                   xcd_remap=remap, xcd_group=G)
        f = k.from_aos(torch.from_numpy(f0).cuda())
        fo, rho, u, _ = k.step(f, 4, mask=mask, tau_field=tau, force=[1e-5, 0.0, 2e-5])
        tg = tau.clone().requires_grad_(True)
        fg = f.clone().requires_grad_(True)
        fN, _, _, _ = rollout(k, fg, 4, tau_field=tg, mask=mask)
        fN.double().pow(2).sum().backward()
        outs.append((fo.view(torch.uint8), rho, u, fg.grad, tg.grad))
    torch.cuda.synchronize()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert torch.equal(a, b)


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (7, 9, 11)), ("D3Q19", (1, 5, 3)),
                                           ("D2Q9", (13, 17)), ("D2Q9", (1, 6))])
def test_link_words_equal_brute_force_link_sets(lattice, shape):
    """lbm_links_build against a NumPy brute force over the mask, word for
    word: bit q when x - e_q (periodic) is a wall, bit 30 when one of those
    walls moves (mask 2), bit 31 alone for a wall cell; second word: the links
    that end at a moving wall."""
    import ctypes
    from maddening_amd import _lib
    from maddening_amd.lbm import _ptr, _stream
    rng = np.random.default_rng(sum(shape))
    m = np.where(rng.random(shape) < 0.2, 1, np.where(rng.random(shape) < 0.15, 2, 0)).astype(np.uint8)
    k = kernel(shape, lattice, 0.8, bb="halfway")
    L = _lib.lib()
    d = k.desc()
    links = torch.empty(L.lbm_links_size(ctypes.byref(d)) // 4, dtype=torch.int32, device="cuda")
    _lib.check(L.lbm_links_build(_stream(k.device), ctypes.byref(d),
                                 _ptr(torch.from_numpy(m).cuda()), _ptr(links)), "links")
    got = links.cpu().numpy().view(np.uint32).reshape((2,) + tuple(shape))
    e, _, _ = orc.lattice(k.Q)
    w0 = np.zeros(shape, np.uint32)
    w1 = np.zeros(shape, np.uint32)
    for q in range(1, k.Q):
        src = np.roll(m, shift=tuple(int(v) for v in e[q]), axis=tuple(range(len(shape))))
        w0 |= np.where(src != 0, np.uint32(1 << q), np.uint32(0))
        w1 |= np.where(src == 2, np.uint32(1 << q), np.uint32(0))
    w0 |= np.where(w1 != 0, np.uint32(1 << 30), np.uint32(0))
    w0 = np.where(m != 0, np.uint32(1 << 31), w0)
    w1 = np.where(m != 0, np.uint32(0), w1)
    assert np.array_equal(got[0], w0)
    assert np.array_equal(got[1], w1)


def test_tune_kernel_picks_a_candidate_and_keeps_results(tmp_path, monkeypatch):
    """The single-lattice tuner times every launch shape, applies the fastest
    and caches it; the tuned kernel gives the same bits as the default one."""
    from maddening_amd import autotune
    monkeypatch.setattr(autotune, "_CACHE", str(tmp_path / "tune.json"))
    shape = (12, 10, 40)
    k = kernel(shape, "D3Q19", 0.8, bb="reference")
    rec = autotune.tune_kernel(k, nsteps=3, reps=1)
    assert len(rec["candidates"]) >= 12
    assert rec["ms_per_step"] == min(c["ms_per_step"] for c in rec["candidates"])
    assert (k.block_x, k.row_vec, k.scheme, k.xcd_remap, k.xcd_group, k.pair) == \
        (rec["block_x"], rec["row_vec"], rec["scheme"], rec["xcd_remap"], rec["xcd_group"],
         rec["pair"])
    assert autotune.tune_kernel(kernel(shape, "D3Q19", 0.8, bb="reference"))["ms_per_step"] == \
        rec["ms_per_step"]  # served from the cache
    rng = np.random.default_rng(0)
    mask = torch.from_numpy((rng.random(shape) < 0.2).astype(np.uint8)).cuda()
    f0 = kernel(shape, "D3Q19", 0.8, bb="reference").equilibrium()
    ref = kernel(shape, "D3Q19", 0.8, bb="reference", scheme="pull").step(f0, 6, mask=mask)[0]
    k.scheme = "pull"
    got = k.step(f0, 6, mask=mask)[0]
    assert torch.equal(got, ref)


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (24, 30, 100)), ("D3Q19", (20, 24, 128)),
                                           ("D2Q9", (300, 1000))])
def test_tune_kernel_times_two_step_launches(tmp_path, monkeypatch, lattice, shape):
    """For a periodic float32 lattice the tuner also times two steps per
    launch; whichever wins, the tuned kernel gives the default kernel's bits."""
    from maddening_amd import autotune
    monkeypatch.setattr(autotune, "_CACHE", str(tmp_path / "tune.json"))
    k = kernel(shape, lattice, 0.8, storage="f32dev")
    rec = autotune.tune_kernel(k, nsteps=4, reps=1)
    assert any(c["pair"] == "force" for c in rec["candidates"])
    assert any(c["pair"] == "off" for c in rec["candidates"])
    assert k.pair == rec["pair"]
    f0 = kernel(shape, lattice, 0.8, storage="f32dev").equilibrium()
    ref = kernel(shape, lattice, 0.8, storage="f32dev", scheme="pull", pair="off").step(f0, 7)[0]
    k.scheme = "pull"
    assert torch.equal(k.step(f0, 7)[0], ref)
