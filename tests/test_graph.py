"""The minimal graph harness (maddening_amd/graph.py), checked the way the
reference checks GraphManager's calling contract: schedule order, forward
vs back edges, additive edges, flux edges, external inputs, and the
equivalence of step / run / run_scan / run_scan_with_history / run_sweep.
The contract itself runs on the CPU with small torch nodes; the LBM node's
fused scan and batched sweep run on the GPU."""
import numpy as np
import pytest
import torch

from maddening_amd.graph import GraphManager
from maddening_amd.node import SimulationNode


class Counter(SimulationNode):
    """x <- x + 1 + inc (inc from boundary inputs)."""

    def initial_state(self):
        return {"x": torch.tensor(float(self.params.get("x0", 0.0)))}

    def update(self, state, bi, dt):
        return {"x": state["x"] + 1.0 + bi.get("inc", 0.0)}


class Doubler(SimulationNode):
    """y <- 2 * src (src from boundary inputs); exposes y + 100 as a flux."""

    def initial_state(self):
        return {"y": torch.tensor(0.0)}

    def update(self, state, bi, dt):
        return {"y": 2.0 * bi.get("src", torch.tensor(0.0))}

    def compute_boundary_fluxes(self, state, bi, dt):
        return {"y_plus": state["y"] + 100.0}


def test_forward_edge_reads_this_step_and_order_is_topological():
    gm = GraphManager()
    gm.add_node(Doubler("d", 1.0))      # inserted first, but depends on c
    gm.add_node(Counter("c", 1.0))
    gm.add_edge("c", "d", "x", "src")
    gm.compile()
    assert gm._schedule == ["c", "d"]
    st = gm.step()
    assert float(st["c"]["x"]) == 1.0
    assert float(st["d"]["y"]) == 2.0       # saw c's new x


def test_back_edge_reads_previous_step():
    gm = GraphManager()
    gm.add_node(Counter("c", 1.0))
    gm.add_node(Doubler("d", 1.0))
    gm.add_edge("c", "d", "x", "src")
    gm.add_edge("d", "c", "y", "inc")       # closes a cycle: back edge
    gm.compile()
    assert gm._schedule == ["c", "d"]
    xs, ys = [], []
    for _ in range(3):
        st = gm.step()
        xs.append(float(st["c"]["x"]))
        ys.append(float(st["d"]["y"]))
    # c: x1 = 0+1+y0(=0) = 1; d: y1 = 2; c: x2 = 1+1+2 = 4; y2 = 8; x3 = 4+1+8 = 13
    assert xs == [1.0, 4.0, 13.0] and ys == [2.0, 8.0, 26.0]


def test_additive_edges_fluxes_and_external_inputs():
    gm = GraphManager()
    gm.add_node(Counter("a", 1.0, x0=1.0))
    gm.add_node(Counter("b", 1.0, x0=10.0))
    gm.add_node(Doubler("d", 1.0))
    gm.add_node(Counter("sink", 1.0))
    gm.add_edge("a", "d", "x", "src", additive=True)
    gm.add_edge("b", "d", "x", "src", additive=True)
    gm.add_edge("d", "sink", "y_plus", "inc", transform=lambda v: v / 2)   # a flux field
    gm.add_external_input("a", "inc")
    st = gm.step({"a": {"inc": torch.tensor(5.0)}})
    assert float(st["a"]["x"]) == 7.0                 # 1 + 1 + 5
    assert float(st["d"]["y"]) == 2.0 * (7.0 + 11.0)
    assert float(st["sink"]["x"]) == 1.0 + (36.0 + 100.0) / 2
    st = gm.step()                                   # declared inputs default to zeros
    assert float(st["a"]["x"]) == 8.0


def test_validation_and_duplicates():
    gm = GraphManager()
    gm.add_node(Counter("c", 1.0))
    with pytest.raises(ValueError, match="already exists"):
        gm.add_node(Counter("c", 1.0))
    gm.add_edge("c", "ghost", "x", "src")
    with pytest.raises(ValueError, match="unknown node 'ghost'"):
        gm.compile()
    gm.remove_node("c")          # drops the edge that references it
    assert gm.validate() == []
    with pytest.raises(KeyError):
        gm.remove_node("c")


def test_run_scan_history_and_callback_agree():
    def build():
        gm = GraphManager()
        gm.add_node(Counter("c", 1.0))
        gm.add_node(Doubler("d", 1.0))
        gm.add_edge("c", "d", "x", "src")
        gm.add_edge("d", "c", "y", "inc")
        return gm
    a, b, c = build(), build(), build()
    seen = []
    a.run(4, callback=lambda i, s: seen.append(float(s["c"]["x"])))
    fb = b.run_scan(4)
    fc, hist = c.run_scan_with_history(4)
    assert float(a.state["c"]["x"]) == float(fb["c"]["x"]) == float(fc["c"]["x"])
    assert hist["c"]["x"].tolist() == seen
    assert hist["d"]["y"].shape == (4,)


def test_run_sweep_batches_initial_states_and_keeps_graph_state():
    gm = GraphManager()
    gm.add_node(Counter("c", 1.0))
    gm.add_external_input("c", "inc")
    before = float(gm.state["c"]["x"])
    init = {"c": {"x": torch.tensor([0.0, 10.0, 20.0])}}
    out = gm.run_sweep(3, init, {"c": {"inc": 1.0}})
    assert out["c"]["x"].tolist() == [6.0, 16.0, 26.0]
    out, hist = gm.run_sweep(2, init, {"c": {"inc": 0.0}}, return_history=True)
    assert hist["c"]["x"].shape == (3, 2)
    assert hist["c"]["x"][1].tolist() == [11.0, 12.0]
    assert float(gm.state["c"]["x"]) == before
    with pytest.raises(ValueError, match="leading batch size"):
        gm.run_sweep(1, {"c": {"x": torch.zeros(2)}, "d": {"y": torch.zeros(3)}})


# --- the LBM node through the harness (GPU) ------------------------------------
@pytest.mark.gpu
def test_lbm_node_scan_equals_steps():
    from maddening_amd.node import LBMNode
    rng = np.random.default_rng(1)
    wall = rng.random((12, 10, 16)) < 0.15

    def build():
        gm = GraphManager()
        gm.add_node(LBMNode("lbm", 1e-3, grid_shape=(12, 10, 16), viscosity=0.08,
                            wall_mask=wall))
        gm.add_external_input("lbm", "body_force", shape=(3,))
        return gm
    ext = {"lbm": {"body_force": [0.0, 1e-5, 2e-5]}}
    a, b = build(), build()
    for _ in range(30):
        a.step(ext)
    fb = b.run_scan(30, ext)                 # one fused launch sequence
    for key in ("f", "density", "velocity", "pressure"):
        torch.testing.assert_close(a.state["lbm"][key], fb["lbm"][key], rtol=2e-6, atol=1e-7)
    _, hist = build().run_scan_with_history(3, ext)
    assert hist["lbm"]["f"].shape == (3, 12, 10, 16, 19)


@pytest.mark.gpu
def test_lbm_sweep_batched_equals_member_loop():
    """The batched sweep (one launch per step for every member) gives each
    member's per-member scan bit for bit."""
    from maddening_amd.node import LBMNode
    rng = np.random.default_rng(2)
    shape = (16, 12, 10)
    gm = GraphManager()
    node = LBMNode("lbm", 1e-3, grid_shape=shape, viscosity=0.1)
    gm.add_node(node)
    B = 3
    k = node.kernel
    fs, masks = [], []
    for b in range(B):
        u = torch.from_numpy(0.02 * rng.standard_normal(shape + (3,))).float().cuda()
        rho = torch.ones(shape, device="cuda")
        fs.append(k.aos_view(k.equilibrium(rho, u)))
        masks.append(torch.from_numpy((rng.random(shape) < 0.1).astype(np.uint8)).cuda())
    init = {"lbm": {"f": torch.stack(fs), "wall_mask": torch.stack(masks)}}
    ext = {"lbm": {"body_force": [1e-5, 0.0, -1e-5]}}
    fast = gm.run_sweep(25, init, ext)
    for b in range(B):
        st = {"f": fs[b], "wall_mask": masks[b]}
        ref = node.run(st, ext["lbm"], 25)
        assert torch.equal(fast["lbm"]["f"][b], ref["f"])
        torch.testing.assert_close(fast["lbm"]["velocity"][b], ref["velocity"], rtol=0, atol=1e-7)
        torch.testing.assert_close(fast["lbm"]["density"][b], ref["density"], rtol=1e-7, atol=0)


@pytest.mark.gpu
def test_lbm_gradient_through_run_scan_matches_finite_differences():
    """torch autograd through GraphManager.run_scan of an LBMNode (the
    counterpart of jax.grad through the reference's run_scan,
    tests/verification/test_gradient_health.py:345-380): dL/dF of a
    velocity loss equals central differences of the float64 forward."""
    from maddening_amd.node import LBMNode
    rng = np.random.default_rng(7)
    wall = rng.random((8, 9, 12)) < 0.15

    def build():
        gm = GraphManager()
        gm.add_node(LBMNode("lbm", 1e-3, grid_shape=(8, 9, 12), viscosity=0.08, wall_mask=wall,
                            storage="f64"))
        gm.add_external_input("lbm", "body_force", shape=(3,), dtype=torch.float64)
        return gm

    ramp = torch.linspace(0, 1, 12, device="cuda", dtype=torch.float64)

    def loss_of(F, grad, kind):
        gm = build()
        with torch.set_grad_enabled(grad):
            st = gm.run_scan(30, {"lbm": {"body_force": F}})["lbm"]
            if kind == "f":      # on the float64 state itself
                return ((st["f"][..., 5] - st["f"][..., 6]) * ramp).mean()
            return (st["velocity"][..., 2].double() * ramp).mean()

    # moments leave the C ABI as float32, so central differences of a
    # velocity loss carry float32 rounding: 1e-4 there, 1e-6 on the f64 state
    # the force enters as float32 (lbm.py:924-931): F and F +- h are chosen
    # float32-exact so that the differences see exactly 2h
    for kind, tol, h in (("f", 1e-6, 2.0 ** -23), ("u", 1e-4, 2.0 ** -20)):
        F = torch.tensor([1e-5, -2e-5, 3e-5], dtype=torch.float32).double().cuda()
        F.requires_grad_(True)
        loss_of(F, True, kind).backward()
        for i in range(3):
            Fp, Fm = F.detach().clone(), F.detach().clone()
            Fp[i] += h
            Fm[i] -= h
            fd = (float(loss_of(Fp, False, kind)) - float(loss_of(Fm, False, kind))) / (2 * h)
            assert abs(float(F.grad[i]) - fd) <= tol * abs(fd) + 1e-12, (kind, i,
                                                                         float(F.grad[i]), fd)
    # gradients with respect to the initial state flow through the harness too
    gm = build()
    st0 = gm.get_node_state("lbm")
    f0 = st0["f"].detach().clone().requires_grad_(True)
    gm.set_node_state("lbm", dict(st0, f=f0))
    st = gm.run_scan(5, {"lbm": {"body_force": [0.0, 0.0, 1e-5]}})
    st["lbm"]["density"].sum().backward()
    # total mass is conserved: d(sum rho_N)/d f0 = 1 for every population
    torch.testing.assert_close(f0.grad, torch.ones_like(f0), rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("n_steps", [10, 100, 1000])
def test_lbm_gradient_health_over_rollout_length(n_steps):
    """Gradient health (tests/verification/test_gradient_health.py:57-75):
    the gradient through a rollout neither vanishes nor blows up as the
    rollout grows."""
    from maddening_amd.node import LBMNode
    node = LBMNode("lbm", 1e-3, grid_shape=(16, 12, 10), viscosity=0.1)
    F = torch.tensor([0.0, 0.0, 1e-5], device="cuda", requires_grad=True)
    st = node.run(node.initial_state(), {"body_force": F}, n_steps)
    st["velocity"][..., 2].mean().backward()
    g = float(F.grad[2])
    assert np.isfinite(g)
    # momentum grows by F per cell and step (periodic, no walls): du/dF ~ n_steps
    assert 0.5 * n_steps < g < 1.5 * n_steps, g
