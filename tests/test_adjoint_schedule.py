"""Checkpoint schedule of the differentiable rollout (maddening_amd.adjoint),
checked on the CPU."""


def test_rollout_auto_checkpoint_choice():
    """checkpoint="auto" keeps every state only when they fit in the stated
    share of free memory (CPU: the choice alone, with free memory given)."""
    from maddening_amd.adjoint import AUTO_MEMORY_SHARE, resolve_checkpoint
    state = 128 * 128 * 256 * 19 * 4          # config-4 state, float32: 318.8 MB
    free = 280 * 2 ** 30                        # an idle 288 GB MI355X
    assert resolve_checkpoint(state, 500, "auto", free_bytes=free) == "all"
    assert resolve_checkpoint(state, 5000, "auto", free_bytes=free) == "sqrt"
    edge = int(AUTO_MEMORY_SHARE * free) // state
    assert resolve_checkpoint(state, edge, "auto", free_bytes=free) == "all"
    assert resolve_checkpoint(state, edge + 1, "auto", free_bytes=free) == "sqrt"
    for other in ("sqrt", "all", 7):
        assert resolve_checkpoint(state, 500, other, free_bytes=0) == other


def _simulate(nsteps, wanted, pairs, monkeypatch):
    """Run adjoint._advance with a recording stand-in for lbm_pull_phase and
    replay its launches symbolically: each state is ('f', 0) for f_0, ('g', k)
    for post-collision g_k, ('F', N) for f_N."""
    from maddening_amd import adjoint

    class Buf:
        n = 0

        def __init__(self):
            Buf.n += 1
            self.id = Buf.n
            self.val = None

    class Kern:
        def empty(self):
            return Buf()

    log = []
    monkeypatch.setattr(adjoint, "_pull", lambda L, st, d, ph, src, dst, *a: log.append((ph, src, dst)))
    f0, fN = Buf(), Buf()
    f0.val = ("f", -1)
    got = {}
    for k, g in adjoint._advance(None, None, None, Kern(), f0, wanted, nsteps, None, None,
                                 None, None, pairs, fN, (None, None, None)):
        got[k] = g
    units = {0: 1, 1: 1, 2: 1, 3: 2, 4: 2, 5: 2, 6: 2}
    for ph, src, dst in log:
        assert src is not dst
        assert src.val is not None, "a launch read a state nothing wrote"
        kind, k = src.val
        assert (kind == "f") == (ph in (0, 4, 6)), (ph, src.val)
        k2 = k + units[ph]
        dst.val = ("F", k2) if ph in (2, 5, 6) else ("g", k2)
    for k, g in got.items():
        assert g.val == ("g", k)          # still holds g_k: not overwritten later
    assert fN.val == ("F", nsteps)
    assert sorted(got) == sorted(wanted)
    return [ph for ph, _, _ in log]


def test_rollout_forward_launch_plan(monkeypatch):
    """The rollout's forward pass keeps exactly the states it is asked for and
    ends with f_N; with two-step launches every run between kept states goes
    two units per launch (phases 3-6), with at most one single launch per run."""
    import math
    for nsteps in (1, 2, 3, 4, 5, 9, 10, 23, 500):
        seg = max(1, math.ceil(math.sqrt(nsteps)))
        s_last = ((nsteps - 1) // seg) * seg
        wanted = sorted({k for k in range(nsteps - 1) if (k + 1) % seg == 0}
                        | set(range(s_last, nsteps - 1)))
        single = _simulate(nsteps, wanted, False, monkeypatch)
        assert len(single) == nsteps + 1 and set(single) <= {0, 1, 2}
        two = _simulate(nsteps, wanted, True, monkeypatch)
        runs = len([k for k in wanted]) + 1
        assert sum(1 for ph in two if ph in (0, 1)) <= runs
        assert len(two) < len(single) or nsteps <= 2
        # all states wanted: nothing to fold but the ends
        every = list(range(nsteps - 1))
        _simulate(nsteps, every, True, monkeypatch)


def test_tune_batch_cache_hit_restores_tiling_and_grid(monkeypatch):
    """A cached batch tuning re-plans the tiles and sets the grid that
    LatticeBatch.run and .graph default to (CPU: a stand-in batch)."""
    import torch
    from maddening_amd import autotune
    from maddening_amd.batch import LatticeBatch

    class FakeBatch:
        kernels, cells, device = [], 123, "cuda:0"
        total_tiles = 10_000
        default_grid = LatticeBatch.default_grid

        def __init__(self):
            self.planned = []
            self.grid = None

        def plan(self, cpt):
            self.planned.append(cpt)
            self.grid = None

    monkeypatch.setattr(torch.cuda, "get_device_name", lambda dev=None: "fake")
    monkeypatch.setattr(autotune, "_load",
                        lambda: _AnyKey({"cells_per_thread": 2, "grid": 4096, "ms_per_launch": 1.0}))
    b = FakeBatch()
    rec = autotune.tune_batch(b)
    assert rec["grid"] == 4096 and b.planned == [2] and b.grid == 4096
    assert b.default_grid() == 4096
    b.total_tiles = 1000
    assert b.default_grid() == 1000       # never more workgroups than tiles


class _AnyKey(dict):
    """A cache stand-in that answers every key with one record."""

    def __init__(self, rec):
        super().__init__()
        self.rec = rec

    def get(self, key, default=None):
        return self.rec
