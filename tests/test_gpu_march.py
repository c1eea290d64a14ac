"""GPU: the general two-step launches against one launch per step.

The two-step march now also (a) collides the stored state in place of a
gather in a sequence's first launch, (b) writes the streamed state and its
moments in the last, so that a single LBMNode.update is one launch, and (c)
takes the reference wall rule (lbm.py:944-948), per-cell force and tau
fields and Zou-He faces (lbm.py:425-510). Each runs the same per-cell
arithmetic as the single-step kernels, so every comparison with one launch
per step is torch.equal; the single update is also held to the float64
oracle at 1e-5 (float32 storage).
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import lbm_numpy as orc

pytestmark = pytest.mark.gpu

# whole rows, 64-column tiles (rows over 341 cells), ragged extents, axes of 1-3 cells
SHAPES = [(9, 20, 128), (5, 7, 64), (3, 4, 342), (4, 3, 13), (2, 1, 700), (6, 6, 6)]


def _inputs(shape, seed, walls=True, fields=True):
    rng = np.random.default_rng(seed)
    mask = (rng.random(shape) < 0.15).astype(np.uint8) if walls else None
    tau = (0.6 + 0.4 * rng.random(shape)).astype(np.float32) if fields else None
    force = (3e-5 * rng.standard_normal(shape + (3,))).astype(np.float32) if fields else None
    rho = 1 + 0.01 * rng.standard_normal(shape)
    u = 0.03 * rng.standard_normal(shape + (3,))
    return mask, tau, force, rho, u


def _run(shape, storage, pair, nsteps, seed, walls=True, fields=True, faces=None,
         scheme="pull"):
    from maddening_amd.lbm import LBMKernel
    mask, tau, force, rho, u = _inputs(shape, seed, walls, fields)
    k = LBMKernel(shape, lattice="D3Q19", tau=0.7, storage=storage, scheme=scheme,
                  bb="reference" if walls else "none", pair=pair)
    f = k.equilibrium(torch.from_numpy(rho).float(), torch.from_numpy(u).float())
    out = k.step(f, nsteps, mask=None if mask is None else torch.from_numpy(mask).cuda(),
                 force=None if force is None else torch.from_numpy(force).cuda(),
                 tau_field=None if tau is None else torch.from_numpy(tau).cuda(), faces=faces)
    torch.cuda.synchronize()
    return k, [t.cpu() for t in out]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("storage", ["f32", "f32dev"])
@pytest.mark.parametrize("nsteps", [1, 2, 3, 4, 7])
def test_general_pairs_equal_single_launches_bitwise(shape, storage, nsteps):
    """Reference walls (15 % of the cells), tau and force fields, an inlet
    and an outlet face on different axes."""
    print(f"general pairs: {shape} {storage} nsteps={nsteps}", flush=True)
    faces = (("y_min", 1.01), ("z_max", 0.995))
    k, forced = _run(shape, storage, "force", nsteps, seed=sum(shape) + nsteps, faces=faces)
    assert k.schedule(nsteps, with_fields=True) == ((nsteps + 2) // 2, (nsteps + 1) // 2)
    k1, single = _run(shape, storage, "off", nsteps, seed=sum(shape) + nsteps, faces=faces)
    assert k1.schedule(nsteps, with_fields=True)[1] == 0
    for a, b in zip(forced, single):
        assert torch.equal(a, b)


@pytest.mark.parametrize("walls,fields", [(True, False), (False, True), (False, False)])
@pytest.mark.parametrize("nsteps", [1, 2, 5])
def test_general_pairs_each_feature_alone(walls, fields, nsteps):
    shape = (7, 11, 96)
    _, forced = _run(shape, "f32dev", "force", nsteps, seed=3, walls=walls, fields=fields)
    _, single = _run(shape, "f32dev", "off", nsteps, seed=3, walls=walls, fields=fields)
    for a, b in zip(forced, single):
        assert torch.equal(a, b)


@pytest.mark.parametrize("storage", ["f16dev", "bf16dev", "f32", "f32dev"])
@pytest.mark.parametrize("walls", [False, True])
def test_one_launch_update_equals_collide_then_stream(storage, walls):
    """A single update by the library's own choice (one launch) against the
    pull sequence's collide and stream + moments launches."""
    shape = (10, 9, 128)
    k, one = _run(shape, storage, "auto", 1, seed=7, walls=walls, scheme="auto")
    assert k.schedule(1, with_fields=True) == (1, 1)
    _, two = _run(shape, storage, "off", 1, seed=7, walls=walls)
    for a, b in zip(one, two):
        assert torch.equal(a, b)


def test_one_launch_update_matches_oracle_f32():
    """LBMNode.update's call, walls + force field + tau field + faces, one
    step from a random state, float32 storage, against the float64 oracle."""
    from maddening_amd.lbm import LBMKernel
    shape = (12, 10, 70)
    mask, tau, force, rho, u = _inputs(shape, 11)
    e, w, _ = orc.lattice(19)
    f0 = orc.equilibrium(rho, u, e, w)
    k = LBMKernel(shape, lattice="D3Q19", tau=0.7, storage="f32", bb="reference")
    assert k.schedule(1) == (1, 1)
    f = k.from_aos(torch.from_numpy(np.ascontiguousarray(f0)).to("cuda", torch.float64))
    faces = (("x_min", 1.01), ("x_max", 0.99))
    fo, r, v, p = k.step(f, 1, mask=torch.from_numpy(mask).cuda(),
                         force=torch.from_numpy(force).cuda(),
                         tau_field=torch.from_numpy(tau).cuda(), faces=faces)
    torch.cuda.synchronize()
    ref = orc.update(f0, tau=tau.astype(np.float64), force=force.astype(np.float64), mask=mask,
                     inlet=("x_min", 1.01 / 3), outlet=("x_max", 0.99 / 3))
    g = k.to_aos(fo, torch.float64).cpu().numpy()
    assert np.max(np.abs(g - ref["f"])) / np.max(np.abs(ref["f"])) < 1e-5
    assert np.max(np.abs(r.cpu().numpy() - ref["density"])) < 1e-5
    assert (np.max(np.abs(v.cpu().numpy() - ref["velocity"]))
            / np.max(np.abs(ref["velocity"]))) < 1e-5
    assert np.all(v.cpu().numpy()[mask != 0] == 0)
    np.testing.assert_allclose(p.cpu().numpy(), r.cpu().numpy() / 3, rtol=1e-6)


@pytest.mark.parametrize("shape", [(9, 20, 128), (3, 5, 700)])
def test_pull_phases_4_5_6(shape):
    """lbm_pull_phase's two-step kinds against the phases they fold:
    4 = 0 then 1, 5 = 1 then 2 (state and moments), 6 = 0 then 2."""
    from maddening_amd import _lib
    from maddening_amd.lbm import LBMKernel, _stream
    mask, tau, force, rho, u = _inputs(shape, 5)
    k = LBMKernel(shape, lattice="D3Q19", tau=0.8, storage="f32dev", bb="reference")
    f = k.equilibrium(torch.from_numpy(rho).float(), torch.from_numpy(u).float())
    m = torch.from_numpy(mask).cuda()
    tf = torch.from_numpy(tau).cuda()
    ff = torch.from_numpy(force).cuda()
    d = k.desc()
    L = _lib.lib()
    s = _stream(k.device)

    def ptr(t):
        return None if t is None else ctypes.c_void_p(t.data_ptr())

    def moments():
        return (torch.empty(shape, device="cuda"), torch.empty(shape + (3,), device="cuda"),
                torch.empty(shape, device="cuda"))

    def phase(ph, src, dst, mom=(None, None, None)):
        _lib.check(L.lbm_pull_phase(s, ctypes.byref(d), ph, ptr(src), ptr(dst), ptr(m), None,
                                    ptr(ff), ptr(tf), *[ptr(t) for t in mom]), "lbm_pull_phase")

    g0, g1, a = k.empty(), k.empty(), k.empty()
    phase(0, f, g0)
    phase(1, g0, g1)
    phase(4, f, a)
    torch.cuda.synchronize()
    assert torch.equal(a.cpu(), g1.cpu())
    fa, fb = k.empty(), k.empty()
    ma, mb = moments(), moments()
    phase(2, g1, fa, ma)
    phase(5, g0, fb, mb)
    torch.cuda.synchronize()
    assert torch.equal(fa.cpu(), fb.cpu())
    for x, y in zip(ma, mb):
        assert torch.equal(x.cpu(), y.cpu())
    fc, mc = k.empty(), moments()
    phase(2, g0, fa, ma)
    phase(6, f, fc, mc)
    torch.cuda.synchronize()
    assert torch.equal(fa.cpu(), fc.cpu())
    for x, y in zip(ma, mc):
        assert torch.equal(x.cpu(), y.cpu())


def test_full_size_update_one_launch():
    """256^3 float32, the node's storage: one update in one launch equals the
    collide + stream launches bit for bit (checksums of all planes and a
    strided sample)."""
    from maddening_amd.lbm import LBMKernel
    shape = (256, 256, 256)
    res = []
    for pair, scheme in (("auto", "auto"), ("off", "pull")):
        k = LBMKernel(shape, lattice="D3Q19", tau=0.8, storage="f32", pair=pair, scheme=scheme)
        rng = np.random.default_rng(1)
        rho = torch.from_numpy(1 + 0.01 * rng.standard_normal(shape).astype(np.float32))
        u = torch.from_numpy(0.03 * rng.standard_normal(shape + (3,)).astype(np.float32))
        f = k.equilibrium(rho, u)
        out, r, v, _ = k.step(f, 1, force=[1e-5, 0.0, 0.0])
        torch.cuda.synchronize()
        res.append((out.view(torch.int32).sum(dtype=torch.int64).item(),
                    out.view(torch.int32)[:, ::17, ::5, ::3].cpu(), r.cpu(), v.cpu()))
        del f, out
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        assert torch.equal(a, b)
