"""GPU: two pull steps per launch (k_pair) against single launches and the oracle.

The two-step launch must give the same bits as the fused single-step
launches it replaces (it runs the same per-cell arithmetic, with the
intermediate state passed through the storage type's encoding exactly as a
store and reload would), so every comparison with single launches is
`torch.equal`. Parity with the float64 oracle is 1e-5 relative for float32
storage, as for the single-step kernels.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import lbm_numpy as orc

pytestmark = pytest.mark.gpu

# whole rows of 1 to 341 cells (1024 / nc rows per workgroup, at least
# three) and longer rows cut into 64-column tiles (the last one partial);
# partial b-tiles (nb not a multiple of the rows per workgroup - 2), partial
# a-chunks, axes of 1-3 cells
SHAPES = [(64, 80, 128), (33, 50, 64), (3, 2, 64), (1, 1, 64), (2, 1, 128), (130, 7, 256),
          (5, 3, 256), (7, 9, 100), (4, 5, 341), (2, 2, 3), (1, 1, 1), (40, 37, 160),
          (3, 4, 342), (5, 14, 512), (2, 1, 1000), (9, 27, 700), (1, 2, 1024)]


def _state(k, shape, seed):
    rng = np.random.default_rng(seed)
    rho = torch.from_numpy(1 + 0.01 * rng.standard_normal(shape)).float()
    u = torch.from_numpy(0.03 * rng.standard_normal(shape + (len(shape),))).float()
    return k.equilibrium(rho, u)


def _run(shape, storage, pair, nsteps, seed=5, force=(1e-5, -2e-5, 3e-5)):
    from maddening_amd.lbm import LBMKernel
    k = LBMKernel(shape, lattice="D3Q19", tau=0.7, storage=storage, scheme="pull", pair=pair)
    f = _state(k, shape, seed)
    out = k.step(f, nsteps, force=list(force))
    torch.cuda.synchronize()
    return k, [t.cpu() for t in out]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("storage", ["f32", "f32dev", "f16dev", "bf16dev"])
@pytest.mark.parametrize("nsteps", [1, 2, 3, 4, 6])
def test_pair_launches_equal_single_launches_bitwise(shape, storage, nsteps):
    print(f"pair vs single: {shape} {storage} nsteps={nsteps}", flush=True)
    k, forced = _run(shape, storage, "force", nsteps)
    launches, pairs = k.schedule(nsteps, force=[1e-5, -2e-5, 3e-5])
    # n + 1 units (collide, n - 1 fused steps, stream + moments), two per launch
    assert pairs == (nsteps + 1) // 2 and launches == (nsteps + 2) // 2
    _, single = _run(shape, storage, "off", nsteps)
    for a, b in zip(forced, single):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(1024, 256, 256), (64, 1024, 1024)])
def test_plain_form_above_4_gib_equals_single_launches_bitwise(shape):
    """The plain form stores through a buffer resource when the output array
    fits a 4 GiB range; 67,108,864 cells of float32 (5.1 GB per state) take
    its other store path: whole rows with one role per wave, and c-tiles."""
    from maddening_amd.lbm import LBMKernel
    print(f"plain pair vs single above 4 GiB: {shape}", flush=True)
    out = {}
    for pair in ("force", "off"):
        k = LBMKernel(shape, lattice="D3Q19", tau=0.7, storage="f32dev", scheme="pull", pair=pair)
        rho = 1 + 0.01 * torch.randn(shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
        u = 0.03 * torch.randn(shape + (3,), device="cuda", generator=torch.Generator("cuda").manual_seed(4))
        f = k.equilibrium(rho, u)
        del rho, u
        out[pair] = k.step(f, 5)
        if pair == "force":
            # (C, SC), (SC, SC), (SC, S)
            assert k.schedule(5) == (3, 3)
        del f, k
    torch.cuda.synchronize()
    for a, b in zip(out["force"], out["off"]):
        assert torch.equal(a, b)
    del out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("shape", [(64, 80, 128), (33, 50, 64), (2, 2, 3), (5, 14, 512)])
@pytest.mark.parametrize("storage", ["f16dev", "bf16dev"])
@pytest.mark.parametrize("nsteps", [2, 5])
def test_16_bit_plain_form_equals_single_launches_bitwise(shape, storage, nsteps):
    """16-bit storage without a force: the plain form of the march, whose
    float16 deviations stay in stored units (2^10 d) through the launch."""
    print(f"16-bit plain pair vs single: {shape} {storage} nsteps={nsteps}", flush=True)
    k, forced = _run(shape, storage, "force", nsteps, force=(0.0, 0.0, 0.0))
    assert k.schedule(nsteps)[1] == (nsteps + 1) // 2
    _, single = _run(shape, storage, "off", nsteps, force=(0.0, 0.0, 0.0))
    for a, b in zip(forced, single):
        assert torch.equal(a, b)


@pytest.mark.parametrize("storage", ["f16dev", "bf16dev"])
def test_16_bit_auto_pairs_equal_single_launches(storage):
    """At 128^3 (the rule's smallest size) 16-bit storage takes two-step
    launches by itself; 10 steps equal one launch per step bit for bit."""
    from maddening_amd.lbm import LBMKernel
    from maddening_amd import workloads
    shape = (128, 128, 128)
    rho, u = workloads.taylor_green_3d(128, 0.05)
    out = {}
    for pair in ("auto", "off"):
        k = LBMKernel(shape, lattice="D3Q19", tau=0.8, storage=storage, pair=pair)
        if pair == "auto":
            # (C, SC, SC) first, then 3 pairs and the last
            assert k.schedule(10) == (5, 5)
        f = k.equilibrium(torch.from_numpy(rho).float(), torch.from_numpy(u).float())
        out[pair] = [t.cpu() for t in k.step(f, 10)]
        torch.cuda.synchronize()
    for a, b in zip(out["auto"], out["off"]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(6, 9, 64), (4, 3, 128)])
def test_pair_launches_match_oracle_f32(shape):
    """Forced two-step launches against the float64 oracle after 7 steps."""
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(21)
    rho0 = 1 + 0.01 * rng.standard_normal(shape)
    u0 = 0.03 * rng.standard_normal(shape + (3,))
    e, w, _ = orc.lattice(19)
    f0 = orc.equilibrium(rho0, u0, e, w)
    F = (1e-5, 0.0, -2e-5)
    k = LBMKernel(shape, lattice="D3Q19", tau=0.8, storage="f32", scheme="pull", pair="force")
    assert k.schedule(7)[1] == 4
    f = k.from_aos(torch.from_numpy(np.ascontiguousarray(f0)).to("cuda", torch.float64))
    f, rho, u, p = k.step(f, 7, force=list(F))
    torch.cuda.synchronize()
    ref = orc.run(f0, 7, tau=0.8, force=F)
    g = k.to_aos(f, torch.float64).cpu().numpy()
    assert np.max(np.abs(g - ref["f"])) / np.max(np.abs(ref["f"])) < 1e-5
    assert np.max(np.abs(u.cpu().numpy() - ref["velocity"])) / np.max(np.abs(ref["velocity"])) < 1e-5
    assert np.max(np.abs(rho.cpu().numpy() - ref["density"])) < 1e-5


@pytest.mark.parametrize("storage", ["f32", "f32dev"])
@pytest.mark.parametrize("shape", [(9, 20, 128), (3, 5, 700)])  # whole rows, 64-column tiles
def test_pull_phase_3_equals_two_phase_1(storage, shape):
    from maddening_amd import _lib
    from maddening_amd.lbm import LBMKernel, _stream
    k = LBMKernel(shape, lattice="D3Q19", tau=0.9, storage=storage, pair="force")
    f = _state(k, shape, 3)
    d = k.desc(force=[2e-5, 0.0, 1e-5])
    L = _lib.lib()
    s = _stream(k.device)
    mid, two, one = k.empty(), k.empty(), k.empty()

    def ptr(t):
        return ctypes.c_void_p(t.data_ptr())

    def phase(ph, src, dst):
        _lib.check(L.lbm_pull_phase(s, ctypes.byref(d), ph, ptr(src), ptr(dst), None, None, None,
                                    None, None, None, None), "lbm_pull_phase")

    phase(1, f, mid)
    phase(1, mid, two)
    phase(3, f, one)
    torch.cuda.synchronize()
    assert torch.equal(one.cpu(), two.cpu())


def test_full_size_auto_pairs_equal_single_launches():
    """256^3 f32dev, the lattice the library picks two-step launches for by
    itself: 11 steps (6 two-step launches, the first colliding the stored
    state and the last writing the moments) equal one launch per step bit
    for bit."""
    from maddening_amd.lbm import LBMKernel
    shape = (256, 256, 256)
    res = []
    for pair in ("auto", "off"):
        k = LBMKernel(shape, lattice="D3Q19", tau=0.8, storage="f32dev", pair=pair)
        launches, pairs = k.schedule(11)
        assert (launches, pairs) == ((6, 6) if pair == "auto" else (12, 0))
        f = _state(k, shape, 9)
        out, rho, u, _ = k.step(f, 11)
        torch.cuda.synchronize()
        res.append((out.view(torch.int32).sum(dtype=torch.int64).item(),
                    out.view(torch.int32)[:, ::17, ::5, ::3].cpu(), rho.cpu(), u.cpu()))
        del f, out
    assert res[0][0] == res[1][0]
    for a, b in zip(res[0][1:], res[1][1:]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(128, 128, 128), (64, 256, 128), (256, 128, 64),
                                   (80, 128, 224), (64, 64, 640), (64, 70, 1000)])
def test_auto_pairs_on_other_shapes_equal_single_launches(shape):
    """Lattices the library runs two steps per launch by itself besides
    256^3 (at least 2^21 cells, a and b of at least 64 cells, whole rows of
    64-256 cells in multiples of 32 or rows of more than 341 cells in
    64-column tiles): 9 steps equal 9 single launches bit for bit."""
    from maddening_amd.lbm import LBMKernel
    outs = []
    for pair in ("auto", "off"):
        k = LBMKernel(shape, lattice="D3Q19", tau=0.75, storage="f32dev", pair=pair)
        assert k.schedule(9)[1] == (5 if pair == "auto" else 0)
        f = _state(k, shape, 12)
        outs.append([t.cpu() for t in k.step(f, 9, force=[1e-5, 0.0, -1e-5])])
        del f
    torch.cuda.synchronize()
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_pair_launches_random_shapes_bitwise():
    """Seeded random extents (rows of 1-1200 cells, whole or in 64-column
    tiles; other axes 1-40), step counts 2-9, both 32-bit storages, random
    uniform force and tau: forced two-step launches equal one launch per
    step bit for bit."""
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(2024)
    for case in range(32):
        nc = int(rng.integers(1, 342)) if case < 24 else int(rng.integers(342, 1201))
        na = int(rng.integers(1, 41))
        nb = int(rng.integers(1, 41))
        while na * nb * nc > 200_000:
            na = max(1, na // 2)
            nb = max(1, nb // 2)
        shape = (na, nb, nc)
        nsteps = int(rng.integers(2, 10))
        storage = ("f32", "f32dev")[case % 2]
        tau = float(rng.uniform(0.52, 1.8))
        force = [float(v) for v in rng.uniform(-3e-5, 3e-5, 3)]
        print(f"random case {case}: {shape} {storage} nsteps={nsteps} tau={tau:.3f}", flush=True)
        outs = []
        for pair in ("force", "off"):
            k = LBMKernel(shape, lattice="D3Q19", tau=tau, storage=storage, scheme="pull",
                          pair=pair)
            f = _state(k, shape, case)
            outs.append([t.cpu() for t in k.step(f, nsteps, force=force)])
        torch.cuda.synchronize()
        for a, b in zip(*outs):
            assert torch.equal(a, b), (case, shape, storage, nsteps)


# D2Q9: rows of 1 to 1920 cells (one or more 960-cell tiles, the last one
# partial), row counts below and above a 32-row chunk
SHAPES_2D = [(64, 1000), (7, 64), (3, 5), (1, 1), (33, 961), (100, 1920), (40, 2000)]


@pytest.mark.parametrize("shape", SHAPES_2D)
@pytest.mark.parametrize("storage", ["f32", "f32dev", "f16dev", "bf16dev"])
@pytest.mark.parametrize("nsteps", [2, 3, 6])
def test_d2q9_pair_launches_equal_single_launches_bitwise(shape, storage, nsteps):
    from maddening_amd.lbm import LBMKernel
    print(f"D2Q9 pair vs single: {shape} {storage} nsteps={nsteps}", flush=True)
    outs = []
    for pair in ("force", "off"):
        k = LBMKernel(shape, lattice="D2Q9", tau=0.7, storage=storage, scheme="pull", pair=pair)
        assert k.schedule(nsteps)[1] == ((nsteps - 1) // 2 if pair == "force" else 0)
        f = _state(k, shape, 4)
        outs.append([t.cpu() for t in k.step(f, nsteps, force=[2e-5, -1e-5])])
    torch.cuda.synchronize()
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_d2q9_pair_launches_match_oracle_f32():
    from maddening_amd.lbm import LBMKernel
    shape = (9, 70)
    rng = np.random.default_rng(8)
    rho0 = 1 + 0.01 * rng.standard_normal(shape)
    u0 = 0.03 * rng.standard_normal(shape + (2,))
    e, w, _ = orc.lattice(9)
    f0 = orc.equilibrium(rho0, u0, e, w)
    F = (1e-5, -2e-5)
    k = LBMKernel(shape, lattice="D2Q9", tau=0.8, storage="f32", scheme="pull", pair="force")
    f = k.from_aos(torch.from_numpy(np.ascontiguousarray(f0)).to("cuda", torch.float64))
    f, rho, u, p = k.step(f, 7, force=list(F))
    torch.cuda.synchronize()
    ref = orc.run(f0, 7, tau=0.8, force=F)
    g = k.to_aos(f, torch.float64).cpu().numpy()
    assert np.max(np.abs(g - ref["f"])) / np.max(np.abs(ref["f"])) < 1e-5
    assert np.max(np.abs(u.cpu().numpy() - ref["velocity"])) / np.max(np.abs(ref["velocity"])) < 1e-5


@pytest.mark.parametrize("shape", [(2048, 2048), (1536, 1536), (1024, 2048)])
def test_d2q9_auto_pairs_equal_single_launches(shape):
    """Periodic D2Q9 lattices of at least 2^21 cells take two-step launches
    by themselves (a partial last 960-column tile at 1536 and 2048 columns);
    9 steps equal one launch per step bit for bit."""
    from maddening_amd.lbm import LBMKernel
    outs = []
    for pair in ("auto", "off"):
        k = LBMKernel(shape, lattice="D2Q9", tau=0.9, storage="f32dev", pair=pair)
        assert k.schedule(9)[1] == (4 if pair == "auto" else 0)
        f = _state(k, shape, 6)
        outs.append([t.cpu() for t in k.step(f, 9, force=[1e-6, 0.0])])
    torch.cuda.synchronize()
    for a, b in zip(*outs):
        assert torch.equal(a, b)


# rows of whole 64-column tiles and ragged ones, b-tiles of 11 rows cut by
# the lattice, chunks of planes cut by it, axes of one cell
TRI_SHAPES = [(9, 20, 128), (7, 9, 100), (40, 37, 256), (5, 14, 512), (3, 5, 700), (1, 1, 64),
              (1, 1, 68), (2, 3, 205), (12, 7, 204), (6, 30, 256), (2, 2, 70), (1, 12, 130),
              (130, 12, 68)]


def _phase(k, phase, src, dst):
    from maddening_amd import _lib
    from maddening_amd.lbm import _ptr, _stream
    d = k.desc()
    _lib.check(_lib.lib().lbm_pull_phase(_stream(k.device), ctypes.byref(d), phase, _ptr(src),
                                         _ptr(dst), None, None, None, None, None, None, None),
               f"lbm_pull_phase {phase}")


@pytest.mark.parametrize("shape", TRI_SHAPES)
@pytest.mark.parametrize("storage", ["f32", "f32dev", "f16dev", "bf16dev"])
def test_three_unit_launch_equals_three_single_launches_bitwise(shape, storage):
    """Phase 7, (C, SC, SC) in one launch, against phase 0 then phase 1 twice."""
    from maddening_amd.lbm import LBMKernel
    print(f"three units vs singles: {shape} {storage}", flush=True)
    k = LBMKernel(shape, lattice="D3Q19", tau=0.7, storage=storage, scheme="pull")
    f0 = _state(k, shape, 11)
    a, b = k.empty(), k.empty()
    _phase(k, 0, f0, a)
    _phase(k, 1, a, b)
    _phase(k, 1, b, a)
    t = k.empty()
    _phase(k, 7, f0, t)
    torch.cuda.synchronize()
    assert torch.equal(a, t)


def test_16_bit_pair_launches_random_shapes_bitwise():
    """Seeded random extents (rows of 1-1200 cells, whole or in 64-column
    tiles; other axes 1-40), step counts 2-9, random tau, float16 and
    bfloat16 storage without a force (the plain form the library picks by
    itself) and with one (the forced general form): two-unit launches equal
    one launch per step bit for bit."""
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(4048)
    for case in range(16):
        nc = int(rng.integers(1, 342)) if case < 12 else int(rng.integers(342, 1201))
        na = int(rng.integers(1, 41))
        nb = int(rng.integers(1, 41))
        while na * nb * nc > 150_000:
            na = max(1, na // 2)
            nb = max(1, nb // 2)
        shape = (na, nb, nc)
        nsteps = int(rng.integers(2, 10))
        storage = ("f16dev", "bf16dev")[case % 2]
        tau = float(rng.uniform(0.52, 1.8))
        force = [0.0, 0.0, 0.0] if case % 4 < 2 else [float(v) for v in rng.uniform(-3e-5, 3e-5, 3)]
        print(f"16-bit random case {case}: {shape} {storage} nsteps={nsteps} tau={tau:.3f} "
              f"force={force}", flush=True)
        outs = []
        for pair in ("force", "off"):
            k = LBMKernel(shape, lattice="D3Q19", tau=tau, storage=storage, scheme="pull",
                          pair=pair)
            f = _state(k, shape, case)
            outs.append([t.cpu() for t in k.step(f, nsteps, force=force)])
        torch.cuda.synchronize()
        for a, b in zip(*outs):
            assert torch.equal(a, b), (case, shape, storage, nsteps)


def test_three_unit_launch_random_shapes_bitwise():
    """Seeded random extents (rows of 64-1100 cells, whole 64-column tiles
    or a partial last one; other axes 1-40), random tau, every 32- and
    16-bit storage: phase 7 equals phase 0 then phase 1 twice bit for bit."""
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(77)
    storages = ("f32", "f32dev", "f16dev", "bf16dev")
    for case in range(24):
        nc = 64 * int(rng.integers(1, 18)) if case % 2 == 0 else int(rng.integers(64, 1101))
        na = int(rng.integers(1, 41))
        nb = int(rng.integers(1, 41))
        while na * nb * nc > 150_000:
            na = max(1, na // 2)
            nb = max(1, nb // 2)
        shape = (na, nb, nc)
        storage = storages[case % 4]
        tau = float(rng.uniform(0.52, 1.8))
        print(f"three units, random case {case}: {shape} {storage} tau={tau:.3f}", flush=True)
        k = LBMKernel(shape, lattice="D3Q19", tau=tau, storage=storage, scheme="pull")
        f0 = _state(k, shape, case)
        a, b = k.empty(), k.empty()
        _phase(k, 0, f0, a)
        _phase(k, 1, a, b)
        _phase(k, 1, b, a)
        t = k.empty()
        _phase(k, 7, f0, t)
        torch.cuda.synchronize()
        assert torch.equal(a, t), (case, shape, storage)


@pytest.mark.parametrize("shape,pair", [((3, 5, 64), "force"), ((4, 23, 192), "force"),
                                        ((64, 128, 256), "auto")])
@pytest.mark.parametrize("nsteps", [4, 6, 20])
@pytest.mark.parametrize("storage", ["f32", "f32dev", "f16dev", "bf16dev"])
def test_even_sequences_with_a_three_unit_launch_equal_single_launches(shape, pair, nsteps,
                                                                        storage):
    print(f"lbm_step with (C, SC, SC): {shape} {pair} {storage} nsteps={nsteps}", flush=True)
    k, multi = _run(shape, storage, pair, nsteps, force=(0.0, 0.0, 0.0))
    # (C, SC, SC), (nsteps - 4) / 2 pairs, (SC, S)
    assert k.schedule(nsteps) == (nsteps // 2, nsteps // 2)
    _, single = _run(shape, storage, "off", nsteps, force=(0.0, 0.0, 0.0))
    for a, b in zip(multi, single):
        assert torch.equal(a, b)
