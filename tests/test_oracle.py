"""Pins of the float64 oracle (oracle/lbm_numpy.py), CPU only.

Each check is chosen so that a plausible slip (dropped term, wrong sign,
wrong index, swapped axis) fails it: closed-form moments, brute-force
permutations, exact conservation laws, analytic flows, and the reference's
own LBM tests (tests/nodes/test_lbm.py) re-run on the oracle.
"""
import itertools

import numpy as np
import pytest

from oracle import lbm_numpy as orc

CS2 = 1.0 / 3.0
LATTICES = [19, 9]


# --- lattice tables ---------------------------------------------------------
@pytest.mark.parametrize("Q", LATTICES)
def test_lattice_tables(Q):
    e, w, opp = orc.lattice(Q)
    D = e.shape[1]
    np.testing.assert_allclose(w.sum(), 1.0, rtol=1e-15)
    assert np.all(opp[opp] == np.arange(Q))
    np.testing.assert_array_equal(e[opp], -e)
    np.testing.assert_array_equal(e[0], np.zeros(D))
    assert w[0] == pytest.approx(1 / 3 if Q == 19 else 4 / 9)
    ef = e.astype(float)
    # isotropy up to fourth order
    np.testing.assert_allclose(np.einsum("q,qa->a", w, ef), 0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, ef, ef), CS2 * np.eye(D), atol=1e-16)
    np.testing.assert_allclose(np.einsum("q,qa,qb,qc->abc", w, ef, ef, ef), 0, atol=1e-16)
    I = np.eye(D)
    iso4 = CS2 ** 2 * (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
                       + np.einsum("ad,bc->abcd", I, I))
    np.testing.assert_allclose(np.einsum("q,qa,qb,qc,qd->abcd", w, ef, ef, ef, ef), iso4,
                               atol=1e-16)


def test_lattice_direction_order_literal():
    """The reference's direction order (lbm.py:92-112, 144-154)."""
    e, _, opp = orc.d3q19()
    assert e[7].tolist() == [1, 1, 0] and e[10].tolist() == [-1, -1, 0]
    assert e[12].tolist() == [-1, 0, 1] and e[17].tolist() == [0, 1, -1]
    assert opp.tolist() == [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
    e2, _, opp2 = orc.d2q9()
    assert e2[6].tolist() == [-1, 1] and e2[7].tolist() == [1, -1]
    assert opp2.tolist() == [0, 2, 1, 4, 3, 8, 7, 6, 5]


def test_product_lattice_matches_oracle():
    from maddening_amd.lattice import d2q9, d3q19
    for lat, (e, w, opp) in ((d3q19(), orc.d3q19()), (d2q9(), orc.d2q9())):
        np.testing.assert_array_equal(lat.e, e)
        np.testing.assert_array_equal(lat.w, w)
        np.testing.assert_array_equal(lat.opp, opp)
        assert lat.cs2 == CS2 and lat.Q == e.shape[0] and lat.D == e.shape[1]


# --- equilibrium / moments / Guo --------------------------------------------
@pytest.mark.parametrize("Q", LATTICES)
def test_equilibrium_moments(Q):
    e, w, _ = orc.lattice(Q)
    D = e.shape[1]
    rng = np.random.default_rng(1)
    rho = 1 + 0.1 * rng.standard_normal((5, 4))
    u = 0.05 * rng.standard_normal((5, 4, D))
    feq = orc.equilibrium(rho, u, e, w)
    ef = e.astype(float)
    np.testing.assert_allclose(feq.sum(-1), rho, rtol=1e-14)
    np.testing.assert_allclose(feq @ ef, rho[..., None] * u, atol=1e-15)
    pi = np.einsum("...q,qa,qb->...ab", feq, ef, ef)
    expect = rho[..., None, None] * (CS2 * np.eye(D) + u[..., :, None] * u[..., None, :])
    np.testing.assert_allclose(pi, expect, atol=1e-14)
    np.testing.assert_allclose(orc.equilibrium(rho, 0 * u, e, w), w * rho[..., None], rtol=1e-15)


@pytest.mark.parametrize("Q", LATTICES)
def test_macroscopic_definition(Q):
    e, _, _ = orc.lattice(Q)
    D = e.shape[1]
    rng = np.random.default_rng(2)
    f = rng.random((3, 4, Q))
    F = rng.standard_normal((3, 4, D))
    rho, u = orc.macroscopic(f, e, F)
    for i, j in itertools.product(range(3), range(4)):
        r = sum(f[i, j, q] for q in range(Q))
        m = [sum(f[i, j, q] * e[q, d] for q in range(Q)) + 0.5 * F[i, j, d] for d in range(D)]
        assert rho[i, j] == pytest.approx(r, rel=1e-14)
        np.testing.assert_allclose(u[i, j], np.array(m) / r, rtol=1e-13)


@pytest.mark.parametrize("Q", LATTICES)
@pytest.mark.parametrize("tau", [0.6, 1.3])
def test_guo_source_moments(Q, tau):
    """sum S = 0, sum S e = (1 - 1/2tau) F, sum S e e = (1 - 1/2tau)(uF + Fu)."""
    e, w, _ = orc.lattice(Q)
    D = e.shape[1]
    rng = np.random.default_rng(3)
    u = 0.05 * rng.standard_normal((6, D))
    F = 1e-3 * rng.standard_normal((6, D))
    S = orc.guo_forcing(u, F, tau, e, w)
    ef = e.astype(float)
    c = 1 - 0.5 / tau
    np.testing.assert_allclose(S.sum(-1), 0, atol=1e-18)
    np.testing.assert_allclose(S @ ef, c * F, rtol=1e-12)
    pi = np.einsum("nq,qa,qb->nab", S, ef, ef)
    np.testing.assert_allclose(pi, c * (u[:, :, None] * F[:, None, :] + F[:, :, None] * u[:, None, :]),
                               atol=1e-17)


# --- streaming / bounce-back -------------------------------------------------
@pytest.mark.parametrize("Q,shape", [(19, (3, 4, 5)), (9, (4, 5))])
def test_stream_is_the_shift_permutation(Q, shape):
    e, _, _ = orc.lattice(Q)
    rng = np.random.default_rng(4)
    f = rng.random(shape + (Q,))
    out = orc.stream(f, e)
    for x in itertools.product(*[range(n) for n in shape]):
        for q in range(Q):
            src = tuple((xi - ei) % n for xi, ei, n in zip(x, e[q], shape))
            assert out[x + (q,)] == f[src + (q,)]


def test_reference_bounce_back_links_brute_force():
    e, w, opp = orc.d3q19()
    rng = np.random.default_rng(5)
    shape = (4, 3, 5)
    f = rng.random(shape + (19,))
    wall = rng.random(shape) < 0.4
    out = orc.bounce_back_reference(orc.stream(f, e), wall, opp)
    for x in itertools.product(*[range(n) for n in shape]):
        for q in range(19):
            if wall[x]:
                # slot q of a wall cell holds what streamed into slot opp(q): it
                # left x + e_q heading opp(q)
                src = tuple((xi + ei) % n for xi, ei, n in zip(x, e[q], shape))
                assert out[x + (q,)] == f[src + (opp[q],)]
            else:
                src = tuple((xi - ei) % n for xi, ei, n in zip(x, e[q], shape))
                assert out[x + (q,)] == f[src + (q,)]


# --- whole-step invariants ----------------------------------------------------
def test_uniform_flow_is_a_fixed_point():
    e, w, _ = orc.d3q19()
    rho = np.full((4, 5, 6), 1.02)
    u = np.broadcast_to(np.array([0.03, -0.01, 0.02]), (4, 5, 6, 3))
    f = orc.equilibrium(rho, u, e, w)
    out = orc.update(f, tau=0.7)
    np.testing.assert_allclose(out["f"], f, rtol=1e-14)


@pytest.mark.parametrize("Q,shape", [(19, (6, 5, 7)), (9, (9, 7))])
def test_mass_and_momentum_growth_are_exact(Q, shape):
    """Periodic: total mass conserved; total raw momentum grows by exactly F per
    cell and step (collision + Guo source + F/2 correction together)."""
    e, w, _ = orc.lattice(Q)
    D = e.shape[1]
    rng = np.random.default_rng(6)
    f = orc.equilibrium(1 + 0.05 * rng.standard_normal(shape),
                        0.03 * rng.standard_normal(shape + (D,)), e, w)
    F = 1e-4 * rng.standard_normal(shape + (D,))
    ef = e.astype(float)
    m0, p0 = f.sum(), (f @ ef).sum(axis=tuple(range(D)))
    tau_field = rng.uniform(0.55, 1.4, shape)
    out = orc.run(f, 5, tau=tau_field, force=F)
    m1, p1 = out["f"].sum(), (out["f"] @ ef).sum(axis=tuple(range(D)))
    assert m1 == pytest.approx(m0, rel=1e-14)
    np.testing.assert_allclose(p1 - p0, 5 * F.sum(axis=tuple(range(D))), atol=1e-13)


def test_taylor_green_viscous_decay_rate():
    """2-D Taylor-Green kinetic energy decays as exp(-4 nu k^2 t), nu = cs2(tau-1/2)."""
    from maddening_amd import workloads
    n, tau, steps = 48, 0.8, 300
    e, w, _ = orc.d2q9()
    rho, u = workloads.taylor_green_2d(n, n, 0.01)
    f = orc.equilibrium(rho, u, e, w)
    out = orc.run(f, steps, tau=tau)
    nu = CS2 * (tau - 0.5)
    k2 = 2 * (2 * np.pi / n) ** 2
    ratio = np.sum(out["velocity"] ** 2) / np.sum(u ** 2)
    assert ratio == pytest.approx(np.exp(-2 * nu * k2 * steps), rel=5e-3)


def test_couette_moving_wall_halfway_exact():
    """Lid (mask 2) moving at U over a static wall with halfway bounce-back:
    the steady profile is linear between the link midpoints y = 0.5 and
    y = ny - 1.5. Pins the sign and size of the moving-wall term."""
    nx, ny, U = 4, 12, 0.05
    mask = np.zeros((nx, ny), np.uint8)
    mask[:, 0] = 1
    mask[:, ny - 1] = 2
    e, w, _ = orc.d2q9()
    f = orc.equilibrium(np.ones((nx, ny)), np.zeros((nx, ny, 2)), e, w)
    out = orc.run(f, 4000, tau=0.9, mask=mask, bb="halfway", lid_velocity=(U, 0.0))
    y = np.arange(1, ny - 1)
    expect = U * (y - 0.5) / (ny - 2)
    np.testing.assert_allclose(out["velocity"][1, 1:ny - 1, 0], expect, atol=2e-5 * U / 0.05)
    np.testing.assert_allclose(out["velocity"][..., 1], 0, atol=1e-10)


def test_halfway_poiseuille_body_force():
    """Force-driven channel with halfway walls at y = 0.5 and ny - 1.5:
    u(y) = F/(2 nu) (y - 0.5)(ny - 1.5 - y). BGK with halfway bounce-back
    reproduces the parabola exactly (no slip) at tau = 1/2 + sqrt(3/16)."""
    nx, ny, F, tau = 3, 14, 1e-6, 0.5 + np.sqrt(3.0 / 16.0)
    mask = np.zeros((nx, ny), np.uint8)
    mask[:, 0] = mask[:, ny - 1] = 1
    e, w, _ = orc.d2q9()
    f = orc.equilibrium(np.ones((nx, ny)), np.zeros((nx, ny, 2)), e, w)
    out = orc.run(f, 8000, tau=tau, mask=mask, bb="halfway", force=(F, 0.0))
    nu = CS2 * (tau - 0.5)
    y = np.arange(1, ny - 1)
    expect = F / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    np.testing.assert_allclose(out["velocity"][1, 1:ny - 1, 0], expect, rtol=1e-8)


def test_reference_bounce_back_poiseuille_amplitude():
    """Reference wall rule, body force, steady channel. Exact pins: the profile
    is a parabola of curvature -F/nu (second differences, nu = cs2(tau-1/2)),
    symmetric, with zero velocity on wall cells. The node-swap rule adds a
    uniform slip, so the amplitude is checked against the reference's own
    Poiseuille bounds (test_lbm_poiseuille.py:150-162: 0.75-1.30 of analytic)."""
    nx, ny, F, tau = 3, 14, 1e-6, 0.8
    mask = np.zeros((nx, ny), np.uint8)
    mask[:, 0] = mask[:, ny - 1] = 1
    e, w, _ = orc.d2q9()
    f = orc.equilibrium(np.ones((nx, ny)), np.zeros((nx, ny, 2)), e, w)
    out = orc.run(f, 8000, tau=tau, mask=mask, force=(F, 0.0))
    nu = CS2 * (tau - 0.5)
    y = np.arange(1, ny - 1)
    prof = out["velocity"][1, 1:ny - 1, 0]
    np.testing.assert_allclose(np.diff(prof, 2), -F / nu, rtol=1e-6)
    np.testing.assert_allclose(prof, prof[::-1], rtol=1e-10)
    expect = F / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    assert 0.75 < prof.max() / expect.max() < 1.30
    assert np.all(out["velocity"][:, [0, ny - 1]] == 0.0)


# --- the reference's own LBM tests, run on the oracle ------------------------
def test_reference_equilibrium_stable_100_steps():
    """test_lbm.py:222-241."""
    st = orc.initial_state((8, 8, 8), 19)
    out = orc.run(st["f"], 100, tau=0.5 + 0.1 / CS2)
    np.testing.assert_allclose(out["density"], 1.0, rtol=1e-3)
    np.testing.assert_allclose(out["velocity"], 0.0, atol=1e-4)


def test_reference_mass_conserved_periodic():
    """test_lbm.py:336-357 (16^3, 50 steps, rtol 1e-5), run in float32."""
    st = orc.initial_state((16, 16, 16), 19, dtype=np.float32)
    m0 = float(np.sum(st["density"]))
    out = orc.run(st["f"], 50, tau=0.5 + 0.1 / CS2)
    assert float(np.sum(out["density"])) == pytest.approx(m0, rel=1e-5)


def test_reference_wall_cells_zero_velocity():
    """test_lbm.py:364-387."""
    wall = np.zeros((8, 8, 8), np.uint8)
    wall[3:5] = 1
    F = np.zeros((8, 8, 8, 3))
    F[..., 0] = 0.001
    st = orc.initial_state((8, 8, 8), 19)
    out = orc.run(st["f"], 20, tau=0.8, mask=wall, force=F)
    np.testing.assert_allclose(out["velocity"][3:5], 0.0, atol=1e-10)


def test_reference_pressure_poiseuille_2d():
    """test_lbm.py:248-329: Zou-He pressure inlet/outlet, 64x16, 2000 steps,
    normalised profile within atol 0.15 of the parabola."""
    nx, ny, nu = 64, 16, 0.1
    wall = np.zeros((nx, ny), np.uint8)
    wall[:, 0] = wall[:, ny - 1] = 1
    st = orc.initial_state((nx, ny), 9, dtype=np.float32)
    out = orc.run(st["f"], 2000, tau=0.5 + nu / CS2, mask=wall,
                  inlet=("x_min", CS2 * 1.01), outlet=("x_max", CS2 * 1.00))
    prof = out["velocity"][nx // 2, :, 0]
    assert abs(prof[0]) < 1e-6 and abs(prof[-1]) < 1e-6
    inner = prof[1:-1]
    assert np.all(inner > 0)
    assert abs(int(np.argmax(inner)) + 1 - ny // 2) <= 2
    H = ny - 2
    y = np.arange(1, ny - 1) - 0.5
    ana = y * (H - y)
    np.testing.assert_allclose(inner / inner.max(), ana / ana.max(), atol=0.15)


def test_oracle_float32_mode_stays_float32():
    st = orc.initial_state((4, 4, 4), 19, dtype=np.float32)
    out = orc.update(st["f"], tau=0.8, force=(1e-5, 0.0, 0.0))
    assert all(out[k].dtype == np.float32 for k in ("f", "density", "velocity", "pressure"))


# --- Zou-He pressure face: closed forms of the reference's rule ---------------
# (lbm.py:425-510). Its normal velocity is u_n = +-(1 - (sum known + sum
# tangential) / rho_p), without the textbook factor 2 on the known sum
# (lbm.py:476-479), applied through f_q = f_opp + feq_q - feq_opp at (rho_p,
# u_n) (:503). Two consequences fix that rule and fail the textbook one:
#  * on a resting face f = w rho_p the unknown weights sum to 1/6 (both
#    lattices), so u_n = +-1/6 and every unknown slot becomes exactly
#    2 w_q rho_p (the textbook u_n is 0 there and leaves w_q rho_p);
#  * for any populations the face cell's density afterwards is
#    rho_p + sum known (the textbook rule gives rho_p).
FACES_OF = {19: ["x_min", "x_max", "y_min", "y_max", "z_min", "z_max"],
            9: ["x_min", "x_max", "y_min", "y_max"]}
ZH_CASES = [(Q, face) for Q in (19, 9) for face in FACES_OF[Q]]


def _face_setup(Q, face):
    e, w, opp = orc.lattice(Q)
    D = e.shape[1]
    shape = (5, 4, 6) if D == 3 else (5, 6)
    axis, side = orc.FACES[face]
    sl = [slice(None)] * D
    sl[axis] = 0 if side == "min" else shape[axis] - 1
    wall = np.zeros(shape, bool)
    idx = [1] * D
    idx[axis] = sl[axis]
    wall[tuple(idx)] = True                     # one wall cell on the face
    wall[(2,) * D] = True                       # and one off it
    return e, w, opp, shape, axis, side, tuple(sl), wall


def _check_resting_face(zh, Q, face, rho_p=1.03):
    e, w, opp, shape, axis, side, sl, wall = _face_setup(Q, face)
    known, unknown, tang = orc.classify_directions(e, axis, side)
    assert w[unknown].sum() == pytest.approx(1 / 6, abs=1e-16)
    f = np.broadcast_to(w * rho_p, shape + (Q,)).copy()
    out = zh(f, rho_p, e, w, face, wall)
    expect = f.copy()
    fluid = ~wall[sl]
    for q in unknown:
        expect[sl + (q,)] = np.where(fluid, 2.0 * w[q] * rho_p, w[q] * rho_p)
    np.testing.assert_allclose(out, expect, rtol=1e-14, atol=0)


def _check_face_density(zh, Q, face, rho_p=0.97):
    e, w, opp, shape, axis, side, sl, wall = _face_setup(Q, face)
    known, unknown, tang = orc.classify_directions(e, axis, side)
    rng = np.random.default_rng(Q + len(face))
    f = w * (1 + 0.2 * rng.standard_normal(shape + (Q,)))
    out = zh(f, rho_p, e, w, face, wall)
    fluid = ~wall[sl]
    rho_new = out[sl].sum(-1)
    expect = rho_p + f[sl][..., known].sum(-1)
    np.testing.assert_allclose(rho_new[fluid], expect[fluid], rtol=1e-13)
    # wall face cells and every cell off the face keep their populations
    np.testing.assert_array_equal(out[sl][~fluid], f[sl][~fluid])
    off = np.ones(shape, bool)
    off[sl] = False
    np.testing.assert_array_equal(out[off], f[off])
    # known and tangential slots are never written
    np.testing.assert_array_equal(out[..., known + tang], f[..., known + tang])


@pytest.mark.parametrize("Q,face", ZH_CASES)
def test_zou_he_resting_face_doubles_unknown_populations(Q, face):
    _check_resting_face(orc.zou_he_pressure_face, Q, face)


@pytest.mark.parametrize("Q,face", ZH_CASES)
def test_zou_he_face_density_is_rho_p_plus_known(Q, face):
    _check_face_density(orc.zou_he_pressure_face, Q, face)


def _textbook_zou_he(f, rho_p, e, w, face, wall):
    """A planted mistake: the textbook Zou-He normal velocity
    u_n = +-(1 - (sum tangential + 2 sum known) / rho_p), otherwise the same
    non-equilibrium bounce-back."""
    axis, side = orc.FACES[face]
    D = e.shape[1]
    known, unknown, tang = orc.classify_directions(e, axis, side)
    sl = [slice(None)] * D
    sl[axis] = 0 if side == "min" else -1
    sl = tuple(sl)
    ff = f[sl]
    s = ff[..., tang].sum(-1) + 2.0 * ff[..., known].sum(-1)
    un = (1.0 - s / rho_p) if side == "min" else (-1.0 + s / rho_p)
    new = ff.copy()
    for q in unknown:
        p = int(orc._opposites(e)[q])
        cand = ff[..., p] + 6.0 * w[q] * rho_p * e[q, axis] * un
        new[..., q] = np.where(wall[sl], ff[..., q], cand)
    out = f.copy()
    out[sl] = new
    return out


@pytest.mark.parametrize("Q,face", [(19, "x_min"), (19, "z_max"), (9, "y_max")])
def test_zou_he_pins_reject_the_textbook_rule(Q, face):
    """Both closed forms fail on the textbook normal velocity, so neither
    can pass on a restatement that slipped back to it."""
    with pytest.raises(AssertionError):
        _check_resting_face(_textbook_zou_he, Q, face)
    with pytest.raises(AssertionError):
        _check_face_density(_textbook_zou_he, Q, face)


@pytest.mark.parametrize("Q", [19, 9])
def test_zou_he_full_update_on_resting_lattice(Q):
    """Through the whole update (lbm.py:899-986): a lattice at rest at
    rho_p with an inlet and an outlet face at that density, one step. Bulk
    cells keep w rho_p; fluid cells of both faces end with 2 w_q rho_p in
    their unknown slots; a wall cell on the inlet face is only bounced,
    which leaves a resting cell unchanged."""
    e, w, opp = orc.lattice(Q)
    D = e.shape[1]
    shape = (6, 5, 4) if D == 3 else (6, 5)
    rho_p = 1.02
    f = np.broadcast_to(w * rho_p, shape + (Q,)).copy()
    mask = np.zeros(shape, np.uint8)
    mask[(0,) + (2,) * (D - 1)] = 1
    out = orc.update(f, tau=0.8, mask=mask, inlet=("x_min", rho_p * CS2),
                     outlet=("x_max", rho_p * CS2))["f"]
    expect = f.copy()
    for side, x in (("min", 0), ("max", shape[0] - 1)):
        _, unknown, _ = orc.classify_directions(e, 0, side)
        for q in unknown:
            expect[x][..., q] = np.where(mask[x] != 0, w[q] * rho_p, 2 * w[q] * rho_p)
    np.testing.assert_allclose(out, expect, rtol=1e-14, atol=0)
