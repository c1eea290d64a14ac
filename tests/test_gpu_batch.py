"""GPU parity of the batched launch (lbm_batch_*): every lattice of a mixed
batch must equal the same lattice stepped alone by lbm_step, bit for bit
(both run the same per-cell update), for every tile size, eagerly and
through a replayed HIP graph."""
import numpy as np
import pytest
import torch

from maddening_amd import workloads

pytestmark = pytest.mark.gpu


def mixed_specs():
    rng = np.random.default_rng(7)
    cav = workloads.cavity_mask(30, 20)
    m2 = (rng.random((33, 17)) < 0.2).astype(np.uint8)
    m3 = (rng.random((8, 9, 20)) < 0.2).astype(np.uint8)
    m4 = workloads.channel_walls((12, 12, 12))
    m4[4:7, 5:8, 3:9] = 2
    return [
        {"lattice": "D2Q9", "shape": (64, 48), "tau": 0.9, "bb": "none"},
        {"lattice": "D2Q9", "shape": (30, 20), "tau": 0.6, "bb": "halfway", "mask": cav,
         "lid_velocity": (0.1, 0.0, 0.0)},
        {"lattice": "D2Q9", "shape": (33, 17), "tau": 1.3, "bb": "reference", "mask": m2,
         "force": (1e-5, -2e-5)},
        {"lattice": "D3Q19", "shape": (16, 12, 10), "tau": 0.8, "bb": "none",
         "force": (0.0, 0.0, 1e-5)},
        {"lattice": "D3Q19", "shape": (8, 9, 20), "tau": 0.55, "bb": "reference", "mask": m3},
        {"lattice": "D3Q19", "shape": (12, 12, 12), "tau": 1.1, "bb": "halfway", "mask": m4,
         "lid_velocity": (0.02, 0.0, -0.01)},
    ]


def initial_states(batch, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for k in batch.kernels:
        rho = torch.from_numpy(1 + 0.01 * rng.standard_normal(k.grid_shape)).float()
        u = torch.from_numpy(0.02 * rng.standard_normal(k.grid_shape + (k.D,))).float()
        out.append(k.equilibrium(rho, u))
    return out


def reference_results(batch, f0s, nsteps):
    res = []
    for k, f0, m, s in zip(batch.kernels, f0s, batch.masks, batch.specs):
        kk = type(k)(k.grid_shape, lattice=k.lat.name, tau=k.tau, storage="f32", bb=k.bb,
                     lid_velocity=k.lid_velocity, scheme="pull", row_vec=-1)
        f, _, _, _ = kk.step(f0, nsteps, mask=m, force=list(s.get("force", (0.0,) * k.D)),
                             moments=False)
        res.append(f)
    return res


@pytest.mark.parametrize("cpt,align,links", [(1, 64, True), (2, 64, True), (4, 64, True),
                                             (1, 1, True), (2, 37, True), (1, 64, False)])
def test_batch_equals_single_lattice_steps(cpt, align, links):
    """Padded planes (align 64 or 37 cells) and the lbm_step layout (align 1),
    with halfway link words or the mask path, give the single-lattice results
    bit for bit."""
    from maddening_amd.batch import LatticeBatch
    batch = LatticeBatch(mixed_specs(), cells_per_thread=cpt, plane_align=align, link_words=links)
    f0s = initial_states(batch)
    nsteps = 11
    ref = reference_results(batch, f0s, nsteps)
    for grid in (7, batch.total_tiles):
        batch.set_states(f0s)
        batch.run(nsteps, grid=grid)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(batch.states(nsteps), ref)):
            assert torch.equal(a, b), (cpt, align, grid, i, float((a - b).abs().max()))


def test_batch_graph_replays_match_eager():
    from maddening_amd.batch import LatticeBatch
    batch = LatticeBatch(mixed_specs(), cells_per_thread=2)
    f0s = initial_states(batch)
    nsteps = 6
    ref = reference_results(batch, f0s, nsteps)
    h = batch.graph(nsteps)
    for replay in range(3):
        batch.set_states(f0s)          # eager work between replays
        torch.cuda.synchronize()
        batch.launch_graph(h)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(batch.states(nsteps), ref)):
            assert torch.equal(a, b), (replay, i)
    batch.close()


def test_per_lattice_baseline_matches():
    from maddening_amd.batch import LatticeBatch
    batch = LatticeBatch(mixed_specs())
    f0s = initial_states(batch)
    batch.set_states(f0s)
    outs = batch.run_per_lattice(5)
    ref = reference_results(batch, f0s, 5)
    for a, b in zip(outs, ref):
        assert torch.equal(a, b)


def test_batch_link_words_follow_mask_updates():
    """Link words are rebuilt from the mask at the start of every run: a
    mask changed in place between runs is seen by the next run."""
    from maddening_amd.batch import LatticeBatch
    batch = LatticeBatch(mixed_specs(), cells_per_thread=1)
    f0s = initial_states(batch)
    nsteps = 4
    batch.set_states(f0s)
    batch.run(nsteps)
    torch.cuda.synchronize()
    i = 5                                   # the D3Q19 halfway lattice
    assert batch.specs[i]["bb"] == "halfway"
    m = batch.masks[i]
    m[5:8, 2:4, 6:10] = 1                   # new static walls, in place
    ref = reference_results(batch, f0s, nsteps)
    batch.set_states(f0s)
    batch.run(nsteps)
    torch.cuda.synchronize()
    assert torch.equal(batch.states(nsteps)[i], ref[i])


def pair_specs():
    """Periodic D3Q19 lattices of several row lengths (two-step class) beside
    lattices that stay one step per launch."""
    return [
        {"lattice": "D3Q19", "shape": (40, 33, 100), "tau": 0.8, "bb": "none",
         "force": (1e-5, 0.0, -1e-5)},
        {"lattice": "D3Q19", "shape": (9, 70, 160), "tau": 0.6, "bb": "none"},
        {"lattice": "D3Q19", "shape": (3, 2, 341), "tau": 1.4, "bb": "none"},
        {"lattice": "D3Q19", "shape": (1, 5, 7), "tau": 0.7, "bb": "none"},
        {"lattice": "D3Q19", "shape": (4, 4, 400), "tau": 0.9, "bb": "none"},   # rows > 341
        {"lattice": "D2Q9", "shape": (64, 48), "tau": 0.9, "bb": "none"},
        {"lattice": "D3Q19", "shape": (8, 9, 20), "tau": 0.55, "bb": "reference",
         "mask": (np.random.default_rng(5).random((8, 9, 20)) < 0.2).astype(np.uint8)},
    ]


@pytest.mark.parametrize("nsteps", [2, 3, 5, 6, 9, 14])
def test_batch_two_step_launches_equal_single_launches(nsteps):
    """The two-step class (an even number of two-step launches, the rest one
    step per launch) equals one launch per step and lbm_step per lattice, bit
    for bit, eagerly and through a replayed graph."""
    from maddening_amd.batch import LatticeBatch
    print(f"batch two-step vs single: nsteps={nsteps}", flush=True)
    on = LatticeBatch(pair_specs(), cells_per_thread=2, pairs=True)
    off = LatticeBatch(pair_specs(), cells_per_thread=2, pairs=False)
    f0s = initial_states(on, seed=11)
    ref = reference_results(on, f0s, nsteps)
    for b in (on, off):
        b.set_states(f0s)
        b.run(nsteps)
    torch.cuda.synchronize()
    for i, (a, c, r) in enumerate(zip(on.states(nsteps), off.states(nsteps), ref)):
        assert torch.equal(a, c), (nsteps, i)
        assert torch.equal(a, r), (nsteps, i)
    h = on.graph(nsteps)
    for _ in range(2):
        on.set_states(f0s)
        on.launch_graph(h)
        torch.cuda.synchronize()
        for i, (a, r) in enumerate(zip(on.states(nsteps), ref)):
            assert torch.equal(a, r), ("graph", nsteps, i)
