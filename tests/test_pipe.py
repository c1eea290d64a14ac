"""LBMPipeNode on the HIP pipe kernels (maddening_amd/pipe.py,
csrc/lbm_pipe.hip) against the NumPy pipe oracle and the reference's own
tests (tests/nodes/test_lbm_pipe.py).

* CPU: constructor validation, masks, initial states (built on the host).
* GPU, f64 storage: every field matches the oracle to 1e-12 (single phase)
  and 1e-11 (multiphase) after 1, 3 and 10 steps.
* GPU, f32 storage: within 1e-5 of the f64 oracle. The reference's
  behavioural tests hold: mass and tracer conservation, propeller drive,
  gravity, stability, bounds, phase separation, settling. run(n) equals n
  updates bit for bit.
"""
import numpy as np
import pytest
import torch

from maddening_amd.pipe import LBMPipeNode
from oracle import lbm_pipe_numpy as op

gpu = pytest.mark.gpu


# --- host side -----------------------------------------------------------------
def test_validation_messages():
    with pytest.raises(ValueError, match="tau must be > 0.5"):
        LBMPipeNode("f", 0.01, tau=0.3, device="cpu")
    with pytest.raises(ValueError, match="tau_tracer must be > 0.5"):
        LBMPipeNode("f", 0.01, tau_tracer=0.4, device="cpu")
    with pytest.raises(ValueError, match="fill_fraction"):
        LBMPipeNode("f", 0.01, fill_fraction=0.0, device="cpu")
    with pytest.raises(ValueError, match="rho_liquid"):
        LBMPipeNode("f", 0.01, nx=8, ny=8, nz=8, G=-5.0, rho_liquid=0.1, rho_gas=0.5,
                    device="cpu")
    with pytest.raises(ValueError, match="rho_gas"):
        LBMPipeNode("f", 0.01, nx=8, ny=8, nz=8, G=-5.0, rho_gas=-0.1, device="cpu")
    assert LBMPipeNode("f", 0.01, tau=0.8, device="cpu").viscosity == pytest.approx(0.1)


def test_masks_match_oracle_and_reference_facts():
    n = LBMPipeNode("f", 0.01, nx=16, ny=16, nz=16, pipe_radius=0.9, propeller_x=5,
                    device="cpu")
    assert np.array_equal(n._wall_mask, op.wall_mask(16, 16, 16, 0.9))
    assert np.array_equal(n._propeller_mask, op.propeller_mask(16, 16, 16, 5, 0.8, 0.9))
    assert n._wall_mask[0, 0, 0] and not n._wall_mask[0, 8, 8]
    assert n._propeller_mask[5].any() and not n._propeller_mask[0].any()
    assert not (n._propeller_mask & n._wall_mask).any()
    p = LBMPipeNode("f", 0.01, nx=16, ny=16, nz=16, fill_fraction=0.55, device="cpu")
    assert not (p._compute_fill_mask() & p._wall_mask).any()


def test_initial_states():
    n = LBMPipeNode("f", 0.01, nx=16, ny=8, nz=8, device="cpu")
    st = n.initial_state()
    assert st["f"].shape == (16, 8, 8, 19) and st["tracer_f"].shape == (16, 8, 8, 7)
    np.testing.assert_allclose(st["density"].numpy(), 1.0, rtol=1e-6)
    np.testing.assert_allclose(st["f"].sum(-1).numpy(), 1.0, rtol=1e-6)
    assert np.all(st["tracer"].numpy()[~n._wall_mask] == 1.0)
    n = LBMPipeNode("f", 0.01, nx=16, ny=16, nz=16, fill_fraction=0.75, device="cpu")
    t = n.initial_state()["tracer"].numpy()[~n._wall_mask]
    assert 0.70 < (t > 0.5).mean() < 0.80
    m = LBMPipeNode("f", 0.01, nx=16, ny=16, nz=16, G=-5.0, fill_fraction=0.55, device="cpu")
    st = m.initial_state()
    rho = st["density"].numpy()[~m._wall_mask]
    tr = st["tracer"].numpy()[~m._wall_mask]
    assert rho.min() < 0.4 and rho.max() > 0.7 and tr.min() < 0.1 and tr.max() > 0.8


def test_run_needs_gpu():
    n = LBMPipeNode("f", 0.01, nx=8, ny=8, nz=8, device="cpu")
    with pytest.raises(RuntimeError, match="GPU"):
        n.update(n.initial_state(), {}, 0.01)


# --- GPU -------------------------------------------------------------------------
def _oracle_kw(node, strength=None):
    return dict(tau=node._tau, tau_tracer=node._tau_tracer, wall=node._wall_mask,
                prop=node._propeller_mask,
                prop_strength=node.params["propeller_strength"] if strength is None else strength,
                gravity=node._gravity, G=node._G, rho_0=node._rho_0, rho_wall=node._rho_wall,
                rho_liquid=node._rho_liquid, rho_gas=node._rho_gas)


def _host(st):
    return {k: v.detach().double().cpu().numpy() for k, v in st.items()}


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


SINGLE = dict(nx=16, ny=12, nz=10, tau=0.8, propeller_x=4, propeller_strength=0.001,
              gravity=-1e-4, fill_fraction=0.75, initial_velocity=0.02, tau_tracer=0.65)
MULTI = dict(nx=16, ny=12, nz=12, tau=0.8, propeller_x=3, propeller_strength=0.01,
             gravity=-1e-3, fill_fraction=0.55, G=-4.5, rho_liquid=1.0, rho_gas=0.25)


@gpu
@pytest.mark.parametrize("cfg,tol", [(SINGLE, 1e-12), (MULTI, 1e-11)], ids=["single", "multi"])
def test_f64_matches_oracle(cfg, tol):
    node = LBMPipeNode("f", 0.01, storage="f64", **cfg)
    st0 = node.initial_state()
    ref = _host(st0)
    kw = _oracle_kw(node)
    for steps in (1, 3, 10):
        out = _host(node.run(st0, {}, steps))
        exp = op.run(ref, steps, **kw)
        for key in ("f", "density", "velocity", "tracer", "tracer_f"):
            assert _rel(out[key], exp[key]) < tol, (steps, key, _rel(out[key], exp[key]))


@gpu
@pytest.mark.parametrize("cfg", [SINGLE, MULTI], ids=["single", "multi"])
def test_f32_close_to_f64_oracle_and_run_equals_updates(cfg):
    node = LBMPipeNode("f", 0.01, **cfg)
    st0 = node.initial_state()
    exp = op.run(_host(st0), 20, **_oracle_kw(node))
    a = node.run(st0, {}, 20)
    out = _host(a)
    for key in ("f", "density", "tracer_f"):
        assert _rel(out[key], exp[key]) < 1e-5, (key, _rel(out[key], exp[key]))
    assert np.max(np.abs(out["velocity"] - exp["velocity"])) < 2e-6
    b = st0
    for _ in range(20):
        b = node.update(b, {}, 0.01)
    for key in a:
        assert torch.equal(a[key], b[key]), key


@gpu
def test_reference_behaviour_single_phase():
    node = LBMPipeNode("fluid", 0.01, nx=16, ny=8, nz=8, propeller_strength=0.0)
    fluid = torch.as_tensor(~node._wall_mask, device="cuda")
    st = node.initial_state()
    m0 = float((st["density"] * fluid).sum())
    t0 = float(st["tracer_f"].sum())
    st = node.run(st, {}, 10)
    assert float((st["density"] * fluid).sum()) == pytest.approx(m0, rel=1e-4)
    st = node.run(st, {}, 10)
    assert float(st["tracer_f"].sum()) == pytest.approx(t0, rel=1e-4)
    # propeller drives +x flow; the boundary input overrides the strength
    for strength_kw, bi in ((dict(propeller_strength=0.001), {}),
                            (dict(propeller_strength=0.0), {"propeller_force": 0.002})):
        n2 = LBMPipeNode("fluid", 0.01, nx=16, ny=8, nz=8, **strength_kw)
        s2 = n2.run(n2.initial_state(), bi, 50)
        assert float((s2["velocity"][..., 0] * fluid).sum() / fluid.sum()) > 0
    # gravity creates z-velocity; 200 steps stay finite with density near 1
    n3 = LBMPipeNode("fluid", 0.01, nx=16, ny=8, nz=8, propeller_strength=0.0,
                     gravity=-0.0005)
    s3 = n3.run(n3.initial_state(), {}, 50)
    assert float((s3["velocity"][..., 2] * fluid).abs().max()) > 0
    n4 = LBMPipeNode("fluid", 0.01, nx=16, ny=8, nz=8, tau=0.8, propeller_strength=0.0005)
    s4 = n4.run(n4.initial_state(), {}, 200)
    assert torch.isfinite(s4["density"]).all() and torch.isfinite(s4["velocity"]).all()
    rho = s4["density"][fluid]
    assert 0.5 < float(rho.min()) and float(rho.max()) < 1.5
    # tracer stays in [0, 1]
    n5 = LBMPipeNode("fluid", 0.01, nx=16, ny=8, nz=8, propeller_strength=0.001,
                     gravity=-0.0001, fill_fraction=0.75)
    s5 = n5.run(n5.initial_state(), {}, 50)
    assert float(s5["tracer"].min()) >= -0.01 and float(s5["tracer"].max()) <= 1.01


@gpu
def test_reference_behaviour_multiphase_through_graph():
    from maddening_amd.graph import GraphManager

    def make(nx=16, prop=0.0, gravity=0.0):
        node = LBMPipeNode("fluid", 0.01, nx=nx, ny=12, nz=12, tau=0.8, pipe_radius=0.9,
                           propeller_x=nx // 6, propeller_radius=0.8, propeller_strength=prop,
                           gravity=gravity, fill_fraction=0.55, G=-4.5, rho_liquid=1.0,
                           rho_gas=0.25, rho_0=1.0)
        gm = GraphManager()
        gm.add_node(node)
        gm.compile()
        return gm, node

    gm, node = make()
    _, hist = gm.run_scan_with_history(200)
    rho = hist["fluid"]["density"].cpu().numpy()
    assert np.isfinite(rho).all() and rho.min() > 0
    fluid = ~node._wall_mask
    c0 = rho[0][fluid].max() - rho[0][fluid].min()
    c1 = rho[99][fluid].max() - rho[99][fluid].min()
    assert c1 >= 0.8 * c0
    gm, node = make(gravity=-0.001)
    _, hist = gm.run_scan_with_history(200)
    t = hist["fluid"]["tracer"][-1].cpu().numpy()
    nz = 12
    assert t[:, :, :nz // 2][fluid[:, :, :nz // 2]].mean() > t[:, :, nz // 2:][fluid[:, :, nz // 2:]].mean()
    still, _ = make(nx=24, prop=0.0, gravity=-0.001)
    moving, _ = make(nx=24, prop=0.01, gravity=-0.001)
    v0 = still.run_scan(200)["fluid"]["velocity"].norm(dim=-1).max()
    v1 = moving.run_scan(200)["fluid"]["velocity"].norm(dim=-1).max()
    assert float(v1) > 1.05 * float(v0)
    n0 = LBMPipeNode("fluid", 0.01, nx=8, ny=8, nz=8, propeller_strength=0.001, gravity=-0.001,
                     G=0.0, fill_fraction=0.55)
    r = n0.run(n0.initial_state(), {}, 20)["density"]
    assert float(r.max()) < 1.1 and float(r.min()) > 0.9
