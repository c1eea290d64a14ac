"""Pins of the pipe-node oracle (oracle/lbm_pipe_numpy.py): the reference's
own helper tests (tests/nodes/test_lbm_pipe.py) restated against the
oracle, plus invariants of the whole step that a dropped term, a wrong sign
or a swapped index would break."""
import numpy as np
import pytest

from oracle import lbm_pipe_numpy as op


def test_constants():
    assert op.E.shape == (19, 3) and op.E7.shape == (7, 3)
    assert abs(op.W.sum() - 1) < 1e-15 and abs(op.W7.sum() - 1) < 1e-15
    for q in range(19):
        assert op.OPP[op.OPP[q]] == q
        assert np.array_equal(op.E[op.OPP[q]], -op.E[q])
    for q in range(7):
        assert op.OPP7[op.OPP7[q]] == q and np.array_equal(op.E7[op.OPP7[q]], -op.E7[q])
    # isotropy: sum w e e = cs2 I for both velocity sets
    for e, w in ((op.E, op.W), (op.E7, op.W7)):
        np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, e, e), np.eye(3) / 3
                                   if e.shape[0] == 19 else np.eye(3) / 4, atol=1e-15)


def test_equilibria():
    rng = np.random.default_rng(0)
    rho = 1 + 0.1 * rng.random((4, 5, 6))
    u = 0.05 * rng.standard_normal((4, 5, 6, 3))
    feq = op.equilibrium(rho, u)
    np.testing.assert_allclose(feq.sum(-1), rho, rtol=1e-14)
    np.testing.assert_allclose(feq @ op.E, rho[..., None] * u, atol=1e-15)
    c = rng.random((4, 5, 6))
    geq = op.tracer_equilibrium(c, u)
    np.testing.assert_allclose(geq.sum(-1), c, rtol=1e-14)
    # D3Q7 first moment: sum g e = C u cs2/cs2_7, cs2_7 = 1/4 -> C u * (1/4)/(1/3)
    np.testing.assert_allclose(geq @ op.E7, c[..., None] * u * 0.75, atol=1e-15)


def test_streaming_conserves_and_shifts():
    rng = np.random.default_rng(1)
    f = rng.random((5, 4, 3, 19))
    s = op.stream(f)
    np.testing.assert_allclose(s.sum(), f.sum(), rtol=1e-14)
    assert s[1, 0, 0, 1] == f[0, 0, 0, 1]          # +x moves one cell along x
    assert s[0, 3, 2, 10] == f[1, 0, 2, 10]        # (-1,-1,0) from (1, 0) wraps to (0, 3)
    g = rng.random((5, 4, 3, 7))
    np.testing.assert_allclose(op.stream(g, op.E7).sum(), g.sum(), rtol=1e-14)


def test_psi_and_eos():
    r = np.linspace(0.01, 5.0, 100)
    assert np.all(np.diff(op.psi(r, 1.0)) > 0)
    assert op.psi(np.array(100.0), 1.0) == pytest.approx(1.0, abs=1e-12)
    assert op.psi(np.array(0.0), 1.0) == 0.0
    np.testing.assert_allclose(op.eos_pressure(np.array([0.5, 1, 2.0]), 0.0, 1.0),
                               np.array([0.5, 1, 2.0]) / 3, rtol=1e-15)


def test_shan_chen_force_properties():
    wall0 = np.zeros((8, 8, 8), bool)
    F = op.shan_chen_force(np.full((8, 8, 8), 0.5), -5.0, 1.0, wall0, 0.5)
    assert np.abs(F).max() < 1e-15
    rho = np.full((16, 4, 4), 0.2)
    rho[8:] = 1.0
    F = op.shan_chen_force(rho, -5.0, 1.0, np.zeros_like(rho, bool), 0.2)
    assert F[7, 2, 2, 0] > 0                    # gas side pulled toward the liquid
    assert abs(F[7, 2, 2, 0]) > 5 * abs(F[2, 2, 2, 0])
    assert abs(F[7, 2, 2, 0]) > 5 * abs(F[13, 2, 2, 0])
    wall = np.zeros((8, 8, 8), bool)
    wall[0] = True
    F = op.shan_chen_force(np.ones((8, 8, 8)), -5.0, 1.0, wall, 1.0)
    assert np.all(F[0] == 0.0)
    # closed form for a 1-D ramp: grad = sum w e psi(x+e) = cs2 * dpsi/dx to second order
    x = np.arange(32.0)
    rho = np.broadcast_to((1.0 + 0.001 * x)[:, None, None], (32, 4, 4)).copy()
    F = op.shan_chen_force(rho, -1.0, 1.0, np.zeros(rho.shape, bool), 1.0)
    p = op.psi(rho, 1.0)
    dpsi = np.exp(-rho)[10, 0, 0] * 0.001
    assert F[10, 0, 0, 0] == pytest.approx(p[10, 0, 0] * dpsi / 3.0, rel=1e-6)


def _state(nx, ny, nz, u0=0.0, seed=None):
    wall = op.wall_mask(nx, ny, nz, 0.9)
    u = np.zeros((nx, ny, nz, 3))
    u[..., 0] = np.where(wall, 0.0, u0)
    rho = np.ones((nx, ny, nz))
    if seed is not None:
        rng = np.random.default_rng(seed)
        rho = rho + 0.01 * rng.random(rho.shape)
        u = u + 0.01 * rng.standard_normal(u.shape) * (~wall)[..., None]
    c = np.where(wall, 0.0, 1.0)
    return wall, {"f": op.equilibrium(rho, u), "density": rho, "velocity": u, "tracer": c,
                  "tracer_f": op.tracer_equilibrium(c, u)}


def test_rest_state_is_fixed_point():
    wall, st = _state(8, 8, 8)
    prop = op.propeller_mask(8, 8, 8, 3, 0.8, 0.9)
    out = op.update(st, tau=0.8, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=0.0)
    np.testing.assert_allclose(out["f"], st["f"], atol=1e-16)
    # a uniform concentration at rest is a fixed point of the tracer too (the
    # pipe's initial tracer, 1 in fluid and 0 in walls, is not: it spreads)
    st["tracer_f"] = op.tracer_equilibrium(np.ones(wall.shape), st["velocity"])
    out = op.update(st, tau=0.8, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=0.0)
    np.testing.assert_allclose(out["tracer_f"], st["tracer_f"], atol=1e-16)


def test_mass_tracer_conservation_and_propeller_drive():
    wall, st = _state(16, 8, 8, seed=3)
    prop = op.propeller_mask(16, 8, 8, 5, 0.8, 0.9)
    kw = dict(tau=0.8, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=0.001,
              gravity=-1e-4)
    m0, t0 = st["f"].sum(), st["tracer_f"].sum()
    out = op.run(st, 30, **kw)
    assert out["f"].sum() == pytest.approx(m0, rel=1e-13)       # bounce-back conserves mass
    assert out["tracer_f"].sum() == pytest.approx(t0, rel=1e-13)
    fluid = ~wall
    assert (out["velocity"][..., 0] * fluid).sum() > 0
    assert np.all(out["velocity"][wall] == 0.0)
    # x-momentum gained equals the propeller impulse on fluid cells, up to wall losses
    assert out["velocity"][..., 0][fluid].mean() < 30 * 0.001


def test_multiphase_branch_without_forces_is_bgk():
    """With no external force and a vanishing G, the exact-difference branch
    must reduce to the single-phase BGK step for the fluid (its forcing
    term is feq(u + F/rho) - feq(u) = 0)."""
    wall, st = _state(8, 8, 8, seed=4)
    prop = op.propeller_mask(8, 8, 8, 2, 0.8, 0.9)
    kw = dict(tau=0.8, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=0.0)
    a = op.run(st, 5, **kw)
    b = op.run(st, 5, G=1e-300, **kw)
    np.testing.assert_allclose(b["f"], a["f"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(b["velocity"], a["velocity"], rtol=0, atol=1e-15)
    with pytest.raises(AssertionError):          # and the forcing term is live
        c = op.run(st, 5, G=-4.5, **kw)
        np.testing.assert_allclose(c["f"], a["f"], rtol=0, atol=1e-6)


def test_multiphase_floor_and_tracer_from_density():
    wall, st = _state(12, 10, 10, seed=5)
    rho = st["density"].copy()
    rho[:, :, 5:] = 0.25
    st["f"] = op.equilibrium(rho, st["velocity"])
    prop = op.propeller_mask(12, 10, 10, 2, 0.8, 0.9)
    out = op.run(st, 20, tau=0.8, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=0.0,
                 gravity=-1e-3, G=-4.5, rho_0=1.0, rho_wall=1.0)
    assert np.all(out["density"] >= 0.01) and np.isfinite(out["f"]).all()
    assert np.array_equal(out["tracer_f"], st["tracer_f"])          # frozen
    t = out["tracer"]
    assert t.min() >= 0 and t.max() <= 1 and np.all(t[wall] == 0)


# --- which velocity each collision uses (lbm_pipe.py:781-787, 833-858) ------
# On a wall-free periodic box in a uniform state at rest, streaming moves
# nothing, so one step of the fluid is exactly its collision. The reference
# collides with the RAW velocity sum f e / rho (velocity_eq = velocity_raw,
# :783) in both the equilibrium and the Guo source, so from rest
#   f_q' = w_q (rho + 3 (1 - 1/(2 tau)) e_q . F)        (S_q at u = 0),
# and the total momentum grows by exactly (1 - 1/(2 tau)) F per cell and step
# (u + F/2 in the equilibrium would make that F). The tracer is collided
# with the step's OUTPUT velocity (:858), which is (1 - 1/(2 tau)) F / rho
# after one step from rest, so its flux after one step is
#   sum_q g_q' e_q = 0.75 C u_new / tau_tracer
# (the pre-step velocity would give 0).
def _uniform_box(n=(4, 3, 5), rho=1.0, c=0.7):
    wall = np.zeros(n, bool)
    prop = np.ones(n, bool)                      # propeller force on every cell
    u = np.zeros(n + (3,))
    st = {"f": op.equilibrium(np.full(n, rho), u), "density": np.full(n, rho), "velocity": u,
          "tracer": np.full(n, c), "tracer_f": op.tracer_equilibrium(np.full(n, c), u)}
    return wall, prop, st


def _check_raw_velocity_collision(update, tau=0.8, Fx=2e-3, gz=-1e-3, rho=1.0):
    wall, prop, st = _uniform_box(rho=rho)
    out = update(st, tau=tau, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=Fx,
                 gravity=gz)
    F = np.array([Fx, 0.0, gz])
    src = 1.0 - 0.5 / tau
    expect = op.W * (rho + 3.0 * src * (op.E @ F))
    np.testing.assert_allclose(out["f"], np.broadcast_to(expect, out["f"].shape),
                               rtol=1e-14, atol=0)
    st3 = op.run(st, 3, tau=tau, tau_tracer=0.6, wall=wall, prop=prop, prop_strength=Fx,
                 gravity=gz) if update is op.update else None
    if st3 is not None:
        mom = (st3["f"] @ op.E).sum(axis=(0, 1, 2))
        np.testing.assert_allclose(mom, 3 * src * F * wall.size, rtol=1e-12, atol=1e-13)


def _check_tracer_velocity(update, tau=0.8, tau_t=0.6, Fx=2e-3, gz=-1e-3, rho=1.0, c=0.7):
    wall, prop, st = _uniform_box(rho=rho, c=c)
    out = update(st, tau=tau, tau_tracer=tau_t, wall=wall, prop=prop, prop_strength=Fx,
                 gravity=gz)
    u_new = (1.0 - 0.5 / tau) * np.array([Fx, 0.0, gz]) / rho
    np.testing.assert_allclose(out["velocity"], np.broadcast_to(u_new, out["velocity"].shape),
                               rtol=1e-13, atol=1e-16)
    flux = out["tracer_f"] @ op.E7
    np.testing.assert_allclose(flux, np.broadcast_to(0.75 * c * u_new / tau_t, flux.shape),
                               rtol=1e-12, atol=1e-19)


def test_pipe_collision_uses_raw_velocity():
    _check_raw_velocity_collision(op.update)


def test_pipe_tracer_collides_with_post_step_velocity():
    _check_tracer_velocity(op.update)


def test_pipe_velocity_pins_reject_the_alternatives(monkeypatch):
    """Planted mistakes in the oracle, each of which one of the two pins
    above must catch: u + F/2 in the fluid equilibrium, u + F/2 in the Guo
    source only, and the tracer collided with the pre-step velocity."""
    orig_eq, orig_guo, orig_teq = op.equilibrium, op.guo_forcing, op.tracer_equilibrium
    # the force of _check_raw_velocity_collision, per cell (rho = 1)
    half = 0.5 * np.array([2e-3, 0.0, -1e-3])

    def shifted_update(patch):
        def upd(st, **kw):
            with monkeypatch.context() as m:
                patch(m)
                return op.update(st, **kw)
        return upd

    def eq_shift(m):
        calls = {"n": 0}

        def eq(rho, u):
            calls["n"] += 1
            return orig_eq(rho, u + half if calls["n"] == 1 else u)
        m.setattr(op, "equilibrium", eq)

    def guo_shift(m):
        m.setattr(op, "guo_forcing", lambda u, F, tau: orig_guo(u + half, F, tau))

    def tracer_pre(m):
        m.setattr(op, "tracer_equilibrium", lambda c, u: orig_teq(c, np.zeros_like(u)))

    for patch in (eq_shift, guo_shift):
        with pytest.raises(AssertionError):
            _check_raw_velocity_collision(shifted_update(patch))
    with pytest.raises(AssertionError):
        _check_tracer_velocity(shifted_update(tracer_pre))
