"""The drop-in LBMNode, checked the way the reference checks its own
(tests/nodes/test_lbm.py): construction and messages, initial state, finite
steps, equilibrium stability, mass conservation, wall velocity, clot
injection, boundary spec, derivatives, fluxes. Validation runs on the CPU;
everything that steps runs on the GPU."""
import numpy as np
import pytest

from maddening_amd.node import LBMNode

gpu = pytest.mark.gpu


# --- construction errors (host only) ------------------------------------------
def test_invalid_lattice():
    with pytest.raises(ValueError, match="Unknown lattice"):
        LBMNode(name="x", timestep=0.001, grid_shape=(8, 8, 8), lattice="D1Q3")


def test_grid_shape_mismatch():
    with pytest.raises(ValueError, match="dims"):
        LBMNode(name="x", timestep=0.001, grid_shape=(8, 8), lattice="D3Q19")


def test_tau_must_exceed_half():
    with pytest.raises(ValueError, match="tau must be > 0.5"):
        LBMNode(name="x", timestep=0.001, grid_shape=(8, 8, 8), viscosity=0.0)


def test_face_validation():
    with pytest.raises(ValueError, match="Unknown inlet_face"):
        LBMNode(name="x", timestep=0.001, grid_shape=(8, 8, 8), inlet_face="w_min")
    with pytest.raises(ValueError, match="only has 2 dimensions"):
        LBMNode(name="x", timestep=0.001, grid_shape=(8, 8), lattice="D2Q9", outlet_face="z_max")


def test_wall_mask_shape_validation():
    with pytest.raises(ValueError, match="wall_mask shape"):
        LBMNode(name="x", timestep=0.001, grid_shape=(8, 8, 8),
                wall_mask=np.zeros((4, 4, 4), dtype=bool))


# --- GPU ---------------------------------------------------------------------
@gpu
def test_default_3d_and_2d():
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(8, 8, 8))
    assert node.name == "lbm" and node._D == 3 and node._Q == 19 and node.tau > 0.5
    node2 = LBMNode(name="lbm2d", timestep=0.001, grid_shape=(16, 8), lattice="D2Q9")
    assert node2._D == 2 and node2._Q == 9
    np.testing.assert_allclose(LBMNode(name="v", timestep=0.001, grid_shape=(8, 8, 8),
                                       viscosity=0.15).tau, 0.5 + 0.15 * 3, rtol=1e-10)
    assert node.halo_width() == {0: 1, 1: 1, 2: 1}
    assert node2.halo_width() == {0: 1, 1: 1}


@gpu
def test_initial_state():
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(8, 6, 4))
    st = node.initial_state()
    assert st["f"].shape == (8, 6, 4, 19)
    assert st["density"].shape == (8, 6, 4)
    assert st["velocity"].shape == (8, 6, 4, 3)
    assert st["pressure"].shape == (8, 6, 4)
    np.testing.assert_allclose(st["density"].cpu(), 1.0, rtol=1e-5)
    np.testing.assert_allclose(st["velocity"].cpu(), 0.0, atol=1e-10)
    np.testing.assert_allclose(st["pressure"].cpu(), 1.0 / 3.0, rtol=1e-5)
    np.testing.assert_allclose(st["f"].sum(-1).cpu(), st["density"].cpu(), rtol=1e-5)
    st2 = LBMNode(name="l2", timestep=0.001, grid_shape=(16, 8), lattice="D2Q9").initial_state()
    assert st2["f"].shape == (16, 8, 9) and st2["velocity"].shape == (16, 8, 2)


@gpu
def test_step_finite_and_equilibrium_stable():
    import torch
    for shape, lat in (((8, 8, 8), "D3Q19"), ((16, 16), "D2Q9")):
        node = LBMNode(name="lbm", timestep=0.001, grid_shape=shape, lattice=lat)
        st = node.initial_state()
        st1 = node.update(st, {}, node.delta_t)
        for key in ("f", "density", "velocity", "pressure"):
            assert torch.isfinite(st1[key]).all()
        for _ in range(100):
            st = node.update(st, {}, node.delta_t)
        np.testing.assert_allclose(st["density"].cpu(), 1.0, rtol=1e-3)
        np.testing.assert_allclose(st["velocity"].cpu(), 0.0, atol=1e-4)


@gpu
def test_mass_conserved_periodic():
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(16, 16, 16), viscosity=0.1)
    st = node.initial_state()
    m0 = float(st["density"].sum())
    for _ in range(50):
        st = node.update(st, {}, node.delta_t)
    assert float(st["density"].sum()) == pytest.approx(m0, rel=1e-5)


@gpu
def test_wall_cells_have_zero_velocity():
    import torch
    wall = np.zeros((8, 8, 8), dtype=bool)
    wall[3:5] = True
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(8, 8, 8), wall_mask=wall)
    st = node.initial_state()
    force = torch.zeros((8, 8, 8, 3), device="cuda")
    force[..., 0] = 0.001
    for _ in range(20):
        st = node.update(st, {"body_force": force}, node.delta_t)
    np.testing.assert_allclose(st["velocity"][3:5].cpu(), 0.0, atol=1e-10)


@gpu
def test_clot_injection_changes_domain():
    import torch
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(16, 8, 8), viscosity=0.1)
    st = node.initial_state()
    force = torch.zeros((16, 8, 8, 3), device="cuda")
    force[..., 0] = 0.0005
    for _ in range(10):
        st = node.update(st, {"body_force": force}, node.delta_t)
    assert float(st["velocity"][8, 3:5, 3:5, 0].abs().mean()) > 1e-6
    clot = np.zeros((16, 8, 8), dtype=bool)
    clot[8, 3:5, 3:5] = True
    for _ in range(10):
        st = node.update(st, {"body_force": force, "wall_mask_update": torch.as_tensor(clot)},
                         node.delta_t)
    np.testing.assert_allclose(st["velocity"][8, 3:5, 3:5].cpu(), 0.0, atol=1e-10)


@gpu
def test_run_equals_repeated_update():
    import torch
    rng = np.random.default_rng(0)
    wall = rng.random((10, 9, 12)) < 0.15
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(10, 9, 12), wall_mask=wall)
    bi = {"body_force": [1e-5, 0.0, 2e-5]}
    a = node.initial_state()
    for _ in range(7):
        a = node.update(a, bi, node.delta_t)
    b = node.run(node.initial_state(), bi, 7)
    for key in ("f", "density", "velocity", "pressure"):
        torch.testing.assert_close(a[key], b[key], rtol=2e-6, atol=1e-7)


@gpu
def test_spec_derivatives_fluxes():
    import torch
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=(8, 8, 8))
    spec = node.boundary_input_spec()
    assert set(spec) == {"inlet_pressure", "outlet_pressure", "body_force", "wall_mask_update"}
    assert spec["body_force"].coupling_type == "additive"
    st = node.initial_state()
    with pytest.raises(NotImplementedError, match="discrete"):
        node.derivatives(st, {})
    fl = node.compute_boundary_fluxes(st, {}, node.delta_t)
    assert torch.isfinite(fl["outlet_pressure_avg"])
    assert float(fl["outlet_pressure_avg"]) == pytest.approx(1.0 / 3.0, rel=1e-6)


@gpu
@pytest.mark.parametrize("face", ["x_min", "x_max", "y_max", "z_min"])
def test_outlet_pressure_average_after_updates(face):
    """compute_boundary_fluxes (lbm.py:995-1017) after eight updates with
    walls and a force: the mean of the float64 oracle's pressure over the
    fluid cells of the outlet face (wall cells left out of sum and count);
    a face that is all wall gives 0 (the count is clamped to 1)."""
    import torch
    from oracle import lbm_numpy as orc
    rng = np.random.default_rng(17)
    shape = (7, 9, 10)
    wall = rng.random(shape) < 0.3
    F = np.zeros(shape + (3,))
    F[..., 0] = 2e-4
    F[..., 1] = -1e-4
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=shape, wall_mask=wall, viscosity=0.08,
                   outlet_face=face)
    st = node.initial_state()
    for _ in range(8):
        st = node.update(st, {"body_force": torch.as_tensor(F, dtype=torch.float32)}, 0.001)
    got = float(node.compute_boundary_fluxes(st, {}, node.delta_t)["outlet_pressure_avg"])
    ref = orc.initial_state(shape, 19, mask=wall.astype(np.uint8))
    out = orc.run(ref["f"], 8, tau=node.tau, force=F.astype(np.float32).astype(np.float64),
                  mask=wall.astype(np.uint8))
    axis, side = {"x_min": (0, 0), "x_max": (0, -1), "y_max": (1, -1), "z_min": (2, 0)}[face]
    p_face = np.take(out["pressure"], side, axis=axis)
    fluid = ~np.take(wall, side, axis=axis)
    want = p_face[fluid].sum() / max(int(fluid.sum()), 1)
    assert fluid.sum() > 0
    assert got == pytest.approx(want, rel=1e-6)
    # the same face with every cell a wall
    wall2 = wall.copy()
    idx = [slice(None)] * 3
    idx[axis] = side
    wall2[tuple(idx)] = True
    node2 = LBMNode(name="lbm", timestep=0.001, grid_shape=shape, wall_mask=wall2, viscosity=0.08,
                    outlet_face=face)
    st2 = node2.update(node2.initial_state(), {}, 0.001)
    assert float(node2.compute_boundary_fluxes(st2, {}, node2.delta_t)["outlet_pressure_avg"]) == 0.0


@gpu
def test_node_matches_oracle_with_walls_and_force():
    """Ten node updates vs the float64 oracle (f32 storage: 1e-5 relative)."""
    import torch
    from oracle import lbm_numpy as orc
    rng = np.random.default_rng(5)
    shape = (9, 10, 11)
    wall = rng.random(shape) < 0.2
    F = np.zeros(shape + (3,))
    F[..., 2] = 1e-4
    node = LBMNode(name="lbm", timestep=0.001, grid_shape=shape, wall_mask=wall, viscosity=0.08)
    st = node.initial_state()
    for _ in range(10):
        st = node.update(st, {"body_force": torch.as_tensor(F, dtype=torch.float32)}, 0.001)
    ref = orc.initial_state(shape, 19, mask=wall.astype(np.uint8))
    out = orc.run(ref["f"], 10, tau=node.tau, force=F.astype(np.float32).astype(np.float64),
                  mask=wall.astype(np.uint8))
    for key in ("f", "density"):
        a = st[key].double().cpu().numpy()
        b = out[key]
        assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-5, key
    # |u| stays below 1e-3 here; storing f ~ w in float32 rounds at ~4e-9 per
    # population and step, so velocity gets the same absolute bound as the
    # kernel tests (1.5e-7)
    err_u = np.max(np.abs(st["velocity"].double().cpu().numpy() - out["velocity"]))
    assert err_u < 1.5e-7, err_u


@gpu
def test_pressure_driven_poiseuille_2d():
    """The reference's own test (tests/nodes/test_lbm.py:248-329) on the GPU node:
    Zou-He inlet/outlet, 64x16 channel, 2000 steps, normalised profile within
    atol 0.15 of the parabola."""
    import torch
    nx, ny = 64, 16
    wall = np.zeros((nx, ny), dtype=bool)
    wall[:, 0] = True
    wall[:, ny - 1] = True
    viscosity = 0.1
    node = LBMNode(name="poiseuille", timestep=0.001, grid_shape=(nx, ny), viscosity=viscosity,
                   lattice="D2Q9", wall_mask=wall, inlet_face="x_min", outlet_face="x_max")
    st = node.initial_state()
    bi = {"inlet_pressure": 1.0 / 3.0 * 1.01, "outlet_pressure": 1.0 / 3.0 * 1.00}
    for _ in range(2000):
        st = node.update(st, bi, node.delta_t)
    assert torch.isfinite(st["velocity"]).all()
    u = st["velocity"][nx // 2, :, 0].double().cpu().numpy()
    np.testing.assert_allclose(u[0], 0.0, atol=1e-6)
    np.testing.assert_allclose(u[ny - 1], 0.0, atol=1e-6)
    inner = u[1:ny - 1]
    assert np.all(inner > 0)
    assert abs(int(np.argmax(inner)) + 1 - ny // 2) <= 2
    H = ny - 2
    y = np.arange(1, ny - 1) - 0.5
    ana = y * (H - y)
    np.testing.assert_allclose(inner / inner.max(), ana / ana.max(), atol=0.15)
    # the same 2000 steps through one multi-step call agree with the updates
    st2 = node.run(node.initial_state(), bi, 2000)
    np.testing.assert_allclose(st2["velocity"].cpu().numpy(), st["velocity"].cpu().numpy(),
                               atol=2e-6)
