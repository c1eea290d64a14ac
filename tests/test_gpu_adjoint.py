"""GPU parity of the hand-written adjoint (lbm_adjoint_step / lbm_moments_adjoint
behind maddening_amd.adjoint.rollout) against

* torch float64 autograd of the oracle restatement (oracle/lbm_torch.py),
  the stand-in for jax.grad of the reference: 1e-10 relative (f64 storage);
* central finite differences of the HIP float64 forward: 1e-6 relative;
* the float64 adjoint, for float32 storage: 1e-4 relative.
"""
import numpy as np
import pytest
import torch

from maddening_amd import workloads
from oracle import lbm_torch as ot

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float((a - b).abs().max() / b.abs().max().clamp(min=1e-300))


def soa_to_aos(t, D):
    return t.permute(*range(1, D + 1), 0)


def f32_exact(x):
    """Round to float32 and back, so both sides see identical inputs."""
    return torch.as_tensor(np.asarray(x, np.float32).astype(np.float64))


def setup(lattice, shape, bb, seed=0):
    rng = np.random.default_rng(seed)
    D = len(shape)
    Q = 19 if D == 3 else 9
    e, w, _, _ = ot.lattice(Q)
    rho = torch.from_numpy(1 + 0.02 * rng.standard_normal(shape))
    u = torch.from_numpy(0.03 * rng.standard_normal(shape + (D,)))
    f0 = ot.equilibrium(rho, u, e, w) + 1e-3 * torch.from_numpy(rng.standard_normal(shape + (Q,)))
    mask = None
    if bb != "none":
        mask = torch.from_numpy((rng.random(shape) < 0.25).astype(np.uint8))
        if bb == "halfway":
            mask[mask.nonzero(as_tuple=True)] = torch.from_numpy(
                rng.integers(1, 3, int(mask.sum())).astype(np.uint8))
    tau = f32_exact(rng.uniform(0.6, 1.3, shape))
    force = f32_exact(1e-4 * rng.standard_normal(shape + (D,)))
    # moment cotangents enter the kernel as float32: make them exact there
    wts = {"f": torch.from_numpy(rng.standard_normal(shape + (Q,))),
           "rho": f32_exact(rng.standard_normal(shape)),
           "u": f32_exact(rng.standard_normal(shape + (D,)))}
    return f0, mask, tau, force, wts


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (5, 6, 7)), ("D2Q9", (7, 9))])
@pytest.mark.parametrize("bb", ["none", "reference", "halfway"])
def test_one_step_vjp_matches_torch_autograd(lattice, shape, bb):
    from maddening_amd.adjoint import rollout
    from maddening_amd.lbm import LBMKernel
    D = len(shape)
    f0, mask, tau, force, wts = setup(lattice, shape, bb)
    lid = (0.05, -0.02, 0.01)[:D] + (0.0,) * (3 - D)
    # oracle
    fo = f0.clone().requires_grad_(True)
    to = tau.clone().requires_grad_(True)
    Fo = force.clone().requires_grad_(True)
    f1, r1, u1 = ot.update(fo, tau=to, force=Fo, mask=mask, bb=bb, lid_velocity=lid)
    Lo = (f1 * wts["f"]).sum() + (r1 * wts["rho"]).sum() + (u1 * wts["u"]).sum()
    Lo.backward()
    # HIP
    k = LBMKernel(shape, lattice=lattice, tau=0.8, storage="f64", bb=bb, lid_velocity=lid)
    fg = k.from_aos(f0.cuda()).requires_grad_(True)
    tg = tau.float().cuda().requires_grad_(True)
    Fg = force.float().cuda().requires_grad_(True)
    f1g, r1g, u1g, _ = rollout(k, fg, 1, force=Fg, tau_field=tg,
                               mask=None if mask is None else mask.cuda(), checkpoint="all")
    Lg = (soa_to_aos(f1g, D).cpu() * wts["f"]).sum() + (r1g.double().cpu() * wts["rho"]).sum() \
        + (u1g.double().cpu() * wts["u"]).sum()
    Lg.backward()
    assert rel(Lg.detach(), Lo.detach()) < 1e-6  # moments are float32 outputs
    assert rel(soa_to_aos(fg.grad, D), fo.grad) < 1e-10
    assert rel(tg.grad, to.grad) < 1e-6   # float32 accumulator
    assert rel(Fg.grad, Fo.grad) < 1e-6   # float32 accumulator


def outlet_uz_aos(f_aos, Fz, mask):
    """Mean u_z over the last plane of axis 2 of an AoS float64 state, with the
    F/2 correction and u = 0 at wall cells (lbm.py:970-976)."""
    f = f_aos[:, :, -1, :]
    e = torch.as_tensor(ot.lattice(19)[3], dtype=torch.float64)
    uz = ((f * e[:, 2]).sum(-1) + 0.5 * Fz) / f.sum(-1)
    return torch.where(mask[:, :, -1] != 0, torch.zeros_like(uz), uz).mean()


@pytest.mark.parametrize("checkpoint", ["sqrt", "all", 4, "auto"])
def test_rollout_gradients_match_torch_autograd(checkpoint):
    from maddening_amd.adjoint import rollout
    from maddening_amd.lbm import LBMKernel
    shape = (8, 7, 12)
    nsteps = 15
    rng = np.random.default_rng(4)
    mask = torch.from_numpy(workloads.sphere_packing(shape, porosity=0.8, rmin=1, rmax=2, seed=1,
                                                     walls=False))
    tau = f32_exact(np.full(shape, 0.7) + 0.05 * rng.standard_normal(shape))
    F = torch.tensor([0.0, 0.0, 2e-5], dtype=torch.float64)
    st = ot.equilibrium(torch.ones(shape, dtype=torch.float64),
                        torch.zeros(shape + (3,), dtype=torch.float64), *ot.lattice(19)[:2])
    # oracle
    to = tau.clone().requires_grad_(True)
    Fo = F.clone().requires_grad_(True)
    f = st
    for _ in range(nsteps):
        f, r, u = ot.update(f, tau=to, force=Fo, mask=mask)
    Lo = outlet_uz_aos(f, Fo[2], mask)
    Lo.backward()
    # HIP
    k = LBMKernel(shape, tau=0.7, storage="f64", bb="reference")
    f0 = k.from_aos(st.cuda())
    tg = tau.float().cuda().requires_grad_(True)
    Fg = F.cuda().requires_grad_(True)
    fN, _, _, _ = rollout(k, f0, nsteps, force=Fg, tau_field=tg, mask=mask.cuda(),
                          checkpoint=checkpoint)
    Lg = outlet_uz_aos(soa_to_aos(fN, 3).cpu(), Fg[2].cpu(), mask)
    Lg.backward()
    assert rel(Lg.detach(), Lo.detach()) < 1e-12
    assert rel(Fg.grad.cpu(), Fo.grad) < 1e-10
    assert rel(tg.grad.cpu(), to.grad) < 1e-5  # float32 accumulator


def outlet_uz_f64(k, f_soa, Fz, mask):
    """Mean u_z over the last plane of axis 2, from a float64 state, with the
    F/2 correction and u = 0 at wall cells (lbm.py:970-976); differentiable."""
    e = torch.as_tensor(k.lat.e, dtype=torch.float64, device=f_soa.device)
    f = f_soa[..., -1]
    rho = f.sum(0)
    uz = ((f * e[:, 2].view(-1, 1, 1)).sum(0) + 0.5 * Fz) / rho
    return torch.where(mask[..., -1] != 0, torch.zeros_like(uz), uz).mean()


def test_rollout_gradient_matches_central_differences():
    """dL/dF and dL/dtau(x) of the HIP float64 adjoint vs central differences of
    the HIP float64 forward (h = 1e-4 relative on F, 1e-4 absolute on tau)."""
    from maddening_amd.adjoint import rollout
    from maddening_amd.lbm import LBMKernel
    shape = (12, 12, 24)
    nsteps = 40
    mask = torch.from_numpy(workloads.sphere_packing(shape, porosity=0.8, rmin=2, rmax=3,
                                                     seed=2)).cuda()
    k = LBMKernel(shape, tau=0.7, storage="f64", bb="reference")
    f0 = k.equilibrium()
    F0 = 1e-5
    tau0 = torch.full(shape, 0.7, device="cuda")

    def loss(Fz, tau_field):
        f, _, _, _ = k.step(f0, nsteps, mask=mask, force=[0.0, 0.0, Fz], tau_field=tau_field,
                            moments=False)
        return float(outlet_uz_f64(k, f, Fz, mask))

    Fg = torch.tensor([0.0, 0.0, F0], dtype=torch.float64, device="cuda", requires_grad=True)
    tg = tau0.clone().requires_grad_(True)
    fN, _, _, _ = rollout(k, f0, nsteps, force=Fg, tau_field=tg, mask=mask)
    outlet_uz_f64(k, fN, Fg[2], mask).backward()
    h = 1e-4 * F0
    fd = (loss(F0 + h, tau0) - loss(F0 - h, tau0)) / (2 * h)
    assert abs(float(Fg.grad[2]) - fd) / abs(fd) < 1e-6
    fluid = (mask == 0).nonzero()
    rng = np.random.default_rng(0)
    gmax = float(tg.grad.abs().max())
    for idx in rng.choice(len(fluid), 4, replace=False):
        c = tuple(int(v) for v in fluid[idx])
        tp, tm = tau0.clone(), tau0.clone()
        tp[c] += 1e-4
        tm[c] -= 1e-4
        # the float32 tau field rounds 0.7 +- 1e-4: divide by the stored step
        hstep = float(tp[c]) - float(tm[c])
        fd_t = (loss(F0, tp) - loss(F0, tm)) / hstep
        g = float(tg.grad[c])
        assert abs(g - fd_t) <= 1e-4 * max(abs(fd_t), 1e-3 * gmax), (c, g, fd_t)


def test_float32_adjoint_close_to_float64():
    from maddening_amd.adjoint import rollout
    from maddening_amd.lbm import LBMKernel
    shape = (10, 10, 16)
    nsteps = 30
    mask = torch.from_numpy(workloads.sphere_packing(shape, porosity=0.8, rmin=2, rmax=3,
                                                     seed=3)).cuda()
    grads = {}
    for storage in ("f32", "f32dev", "f64"):
        k = LBMKernel(shape, tau=0.7, storage=storage, bb="reference")
        Fg = torch.tensor([0.0, 0.0, 1e-5], dtype=torch.float64, device="cuda",
                          requires_grad=True)
        tg = torch.full(shape, 0.7, device="cuda", requires_grad=True)
        _, _, u, _ = rollout(k, k.equilibrium(), nsteps, force=Fg, tau_field=tg, mask=mask)
        u[..., -1, 2].double().mean().backward()
        grads[storage] = (Fg.grad.clone(), tg.grad.clone())
    for s in ("f32", "f32dev"):
        assert rel(grads[s][0], grads["f64"][0]) < 1e-4
        assert rel(grads[s][1], grads["f64"][1]) < 1e-3


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (6, 7, 9)), ("D2Q9", (9, 11))])
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_exchanged_slot_cotangents_match_plain_chain(lattice, shape, storage):
    """Under the reference wall rule, a chain moments-adjoint -> step -> step
    run with cotangents in the exchanged-slot form (LBM_FLAG_ADJ_*_SWAPPED)
    gives the plain chain's dL/df bit for bit once the wall slots are
    exchanged back, and the same dL/dtau."""
    import ctypes
    from maddening_amd import _lib
    from maddening_amd.lbm import LBMKernel, _ptr, _stream
    D = len(shape)
    Q = 19 if D == 3 else 9
    f0, mask, tau, force, wts = setup(lattice, shape, "reference", seed=3)
    k = LBMKernel(shape, lattice=lattice, tau=0.8, storage=storage, bb="reference")
    L = _lib.lib()
    st = _stream(k.device)
    m = mask.cuda()
    tf = tau.float().cuda()
    f = k.from_aos(f0.to(k.device, k.dtype))
    g1 = k.empty()
    _lib.check(L.lbm_pull_phase(st, ctypes.byref(k.desc()), 0, _ptr(f), _ptr(g1), _ptr(m), None,
                                None, _ptr(tf), None, None, None), "pull")
    rho_bar = wts["rho"].float().cuda()
    fbar_N = k.from_aos(wts["f"].to(k.device, k.dtype))
    opp = torch.as_tensor(k.lat.opp, device=k.device).long()
    wall = m.bool()

    def chain(swapped):
        def desc(bits):
            d = k.desc(force=[1e-4, -2e-4, 3e-4][:D])
            if swapped:
                d.flags |= bits
            return d
        IN, OUT = _lib.LBM_FLAG_ADJ_IN_SWAPPED, _lib.LBM_FLAG_ADJ_OUT_SWAPPED
        work = torch.empty(L.lbm_adjoint_workspace_size(ctypes.byref(desc(0))),
                           dtype=torch.uint8, device=k.device)
        _lib.check(L.lbm_adjoint_begin(st, ctypes.byref(desc(0)), _ptr(work)), "begin")
        a, b, c = (torch.empty(k.state_shape, dtype=k.dtype, device=k.device) for _ in range(3))
        tb = torch.zeros(shape, device=k.device)
        _lib.check(L.lbm_moments_adjoint(st, ctypes.byref(desc(OUT)), _ptr(g1), _ptr(m), None,
                                         _ptr(rho_bar), None, None, _ptr(fbar_N), _ptr(a), None,
                                         _ptr(work)), "moments adjoint")
        # step 1 started from f_1 = S B g_1 ... here g_1 stands in for the
        # previous post-collision state; step 0 from f_0 itself
        _lib.check(L.lbm_adjoint_step(st, ctypes.byref(desc(IN | OUT)), _ptr(g1), 1, _ptr(a),
                                      _ptr(b), _ptr(m), None, None, _ptr(tf), _ptr(tb), None,
                                      _ptr(work)), "step 1")
        mid = b.clone()
        _lib.check(L.lbm_adjoint_step(st, ctypes.byref(desc(IN)), _ptr(f), 0, _ptr(b), _ptr(c),
                                      _ptr(m), None, None, _ptr(tf), _ptr(tb), None, _ptr(work)),
                   "step 0")
        out = torch.empty(3, dtype=torch.float64, device=k.device)
        _lib.check(L.lbm_adjoint_end(st, ctypes.byref(desc(0)), _ptr(work), _ptr(out)), "end")
        torch.cuda.synchronize()
        return mid, c, tb, out

    mid_p, out_p, tb_p, F_p = chain(False)
    mid_s, out_s, tb_s, F_s = chain(True)
    # the exchanged form differs from the plain one exactly at wall cells
    mid_back = torch.where(wall.unsqueeze(0), mid_s[opp], mid_s)
    assert torch.equal(mid_back, mid_p)
    assert not torch.equal(mid_s, mid_p)
    assert torch.equal(out_s, out_p)
    assert torch.equal(tb_s, tb_p)
    assert torch.equal(F_s, F_p)


FACE_CASES = [("D2Q9", (10, 7), (("x_min", 1.03), ("x_max", 0.99)), "reference"),
              ("D3Q19", (6, 7, 8), (("z_min", 1.02), ("y_max", 0.98)), "reference"),
              ("D3Q19", (5, 6, 7), (None, ("x_max", 0.97)), "none")]


@pytest.mark.parametrize("lattice,shape,faces,bb", FACE_CASES)
@pytest.mark.parametrize("nsteps,checkpoint", [(1, "all"), (6, "sqrt"), (6, "all")])
def test_adjoint_through_zou_he_faces_matches_torch_autograd(lattice, shape, faces, bb, nsteps,
                                                             checkpoint):
    """Rollouts with Zou-He faces (pressure = rho cs2 on two faces, a shared
    edge in 3-D, or one face only) against autograd of the torch
    restatement: dL/df0, dL/dF and dL/drho of each face, f64 storage."""
    from maddening_amd.adjoint import rollout
    from maddening_amd.lbm import LBMKernel
    D = len(shape)
    f0, mask, _, force, wts = setup(lattice, shape, bb, seed=21)
    Fu = force.reshape(-1, D).mean(0)          # a uniform force
    # oracle
    fo = f0.clone().requires_grad_(True)
    Fo = Fu.clone().requires_grad_(True)
    ro = [None if fc is None else torch.tensor(fc[1], dtype=torch.float64, requires_grad=True)
          for fc in faces]
    f = fo
    for _ in range(nsteps):
        f, r, u = ot.update(f, tau=0.8, force=Fo, mask=mask, bb=bb,
                            inlet=None if faces[0] is None else (faces[0][0], ro[0]),
                            outlet=None if faces[1] is None else (faces[1][0], ro[1]))
    Lo = (f * wts["f"]).sum() + (r * wts["rho"]).sum() + (u * wts["u"]).sum()
    Lo.backward()
    # HIP
    k = LBMKernel(shape, lattice=lattice, tau=0.8, storage="f64", bb=bb)
    fg = k.from_aos(f0.cuda()).requires_grad_(True)
    Fg = Fu.clone().cuda().requires_grad_(True)
    rg = [None if fc is None else torch.tensor(fc[1], dtype=torch.float64, requires_grad=True)
          for fc in faces]
    gf = tuple(None if fc is None else (fc[0], rgi) for fc, rgi in zip(faces, rg))
    f1, r1, u1, _ = rollout(k, fg, nsteps, force=Fg, mask=None if mask is None else mask.cuda(),
                            checkpoint=checkpoint, faces=gf)
    Lg = (soa_to_aos(f1, D).cpu() * wts["f"]).sum() + (r1.double().cpu() * wts["rho"]).sum() \
        + (u1.double().cpu() * wts["u"]).sum()
    Lg.backward()
    assert rel(Lg.detach(), Lo.detach()) < 1e-6
    assert rel(soa_to_aos(fg.grad, D), fo.grad) < 1e-10
    assert rel(Fg.grad.cpu(), Fo.grad) < 1e-9
    for a, b in zip(rg, ro):
        if a is not None:
            assert rel(a.grad, b.grad) < 1e-9, (float(a.grad), float(b.grad))


def test_node_pressure_gradient_matches_finite_differences():
    """d(mean outlet-region velocity)/d(inlet pressure) through the GPU node
    (reference Poiseuille channel, lbm.py:950-968) equals central differences."""
    from maddening_amd.node import LBMNode
    nx, ny = 32, 12
    wall = np.zeros((nx, ny), dtype=bool)
    wall[:, 0] = wall[:, -1] = True
    node = LBMNode("ch", 1e-3, grid_shape=(nx, ny), viscosity=0.1, lattice="D2Q9",
                   wall_mask=wall, inlet_face="x_min", outlet_face="x_max", storage="f64")

    def loss(p_in, grad):
        with torch.set_grad_enabled(grad):
            st = node.run(node.initial_state(), {"inlet_pressure": p_in,
                                                 "outlet_pressure": 1.0 / 3.0}, 200)
            return st["f"][nx // 2, :, 1].sum() - st["f"][nx // 2, :, 2].sum()   # x-momentum

    p = torch.tensor(1.0 / 3.0 * 1.01, dtype=torch.float64, requires_grad=True)
    loss(p, True).backward()
    h = 1e-6
    fd = (float(loss(float(p.detach()) + h, False)) - float(loss(float(p.detach()) - h, False))) \
        / (2 * h)
    assert abs(float(p.grad) - fd) < 1e-6 * abs(fd), (float(p.grad), fd)


@pytest.mark.parametrize("lattice,shape,storage", [("D3Q19", (9, 8, 11), "f32"),
                                                   ("D3Q19", (6, 7, 64), "f64"),
                                                   ("D2Q9", (21, 33), "f32")])
def test_halfway_rollout_link_words_match_mask_path(monkeypatch, lattice, shape, storage):
    """Under halfway bounce-back the rollout builds link words once
    (lbm_links_build) and every forward, recompute and transposed launch reads
    them instead of the neighbours' mask bytes; with the mask path instead,
    the states and all gradients are the same bits."""
    from maddening_amd import adjoint as adj
    from maddening_amd.lbm import LBMKernel
    D = len(shape)
    f0, mask, tau, force, wts = setup(lattice, shape, "halfway", seed=2)
    lid = (0.04, -0.03, 0.02)[:D] + (0.0,) * (3 - D)
    k = LBMKernel(shape, lattice=lattice, tau=0.8, storage=storage, bb="halfway",
                  lid_velocity=lid)
    wf = k.from_aos(wts["f"].cuda()).to(k.dtype)

    def run():
        fg = k.from_aos(f0.cuda().to(k.dtype)).requires_grad_(True)
        tg = tau.float().cuda().requires_grad_(True)
        Fg = torch.tensor([1e-5, -2e-5, 3e-5][:D], dtype=torch.float64, device="cuda",
                          requires_grad=True)
        fN, rho, u, _ = adj.rollout(k, fg, 9, force=Fg, tau_field=tg, mask=mask.cuda(),
                                    checkpoint=3)
        ((fN * wf).sum() + rho.double().sum()).backward()
        torch.cuda.synchronize()
        return fN.detach(), fg.grad, tg.grad, Fg.grad

    with_links = run()
    monkeypatch.setattr(adj, "_links", lambda *a, **kw: None)
    with_mask = run()
    for a, b in zip(with_links, with_mask):
        assert torch.equal(a, b)
