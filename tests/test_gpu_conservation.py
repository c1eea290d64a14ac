"""The collision's conservation step, cell by cell, for every float32-compute
storage type.

In exact arithmetic a BGK collision with the Guo source keeps sum f and
moves sum f e from j to j + F. Rounding each stored population misses both
targets by the sum of 19 roundings. The kernels instead take every stored
population but four as rounded, solve the directions along +a, +b, +c from
the momentum targets and the rest population from the mass target, and
round those as stored, so the stored values miss each target by at most
half an ulp (in the storage format) of the one population solved for it,
plus float32 arithmetic on the targets.

One collide launch (lbm_pull_phase 0: no streaming, so cell x of the output
is cell x of the input after its collision) on a random near-equilibrium
state with a uniform force; the float64 sums of the stored values per cell
are compared with the targets computed from the stored input. A collision
that rounds each population on its own misses the 16-bit targets by far
more than these bounds (tests/test_quantization.py shows it on the float64
oracle's collision).
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import lbm_numpy as orc

pytestmark = pytest.mark.gpu

F = (2e-5, -1e-5, 3e-5)
SLACK = 4e-8  # float32 arithmetic of the targets (|d| < 0.05, 19-term sums)


def _ulp(storage, v):
    """ulp of the stored value v (a deviation) in the storage format."""
    v = np.abs(np.asarray(v, np.float64))
    if storage == "f16dev":
        return np.spacing(np.float16(v * 1024.0)).astype(np.float64) / 1024.0
    if storage == "bf16dev":
        e = np.floor(np.log2(np.maximum(v, 2.0 ** -126)))
        return 2.0 ** (e - 7)
    if storage == "f32":
        return np.zeros_like(v)  # rounding of w + d: in the slack below
    return np.zeros_like(v)


def _deviations(k, f, w):
    """Stored state -> float64 deviations (Q, cells)."""
    a = f.double().cpu().numpy().reshape(k.Q, -1)
    if k.storage == "f16dev":
        return a / 1024.0
    if k.storage == "f32":
        return a - w[:, None]
    return a


@pytest.mark.parametrize("storage", ["f32", "f32dev", "f16dev", "bf16dev"])
def test_collision_meets_mass_and_momentum_targets_per_cell(storage):
    from maddening_amd import _lib
    from maddening_amd.lbm import LBMKernel, _ptr, _stream
    shape = (12, 10, 16)
    rng = np.random.default_rng(7)
    e, w, _ = orc.lattice(19)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = 0.04 * rng.standard_normal(shape + (3,))
    f_aos = orc.equilibrium(rho, u, e, w) * (1.0 + 0.02 * rng.standard_normal(shape + (19,)))
    k = LBMKernel(shape, tau=0.7, storage=storage)
    f0 = k.from_aos(torch.from_numpy(f_aos).cuda())
    g = k.empty()
    d = k.desc(force=F)
    _lib.check(_lib.lib().lbm_pull_phase(_stream(k.device), ctypes.byref(d), 0, _ptr(f0), _ptr(g),
                                         None, None, None, None, None, None, None),
               "lbm_pull_phase")
    torch.cuda.synchronize()
    d0 = _deviations(k, f0, w)
    d1 = _deviations(k, g, w)
    ev = e.astype(np.float64)
    dm = d1.sum(0) - d0.sum(0)
    dj = ev.T @ d1 - (ev.T @ d0 + np.asarray(F)[:, None])
    # the solved populations: rest (mass), +x, +y, +z (momentum)
    qa = [int(np.flatnonzero((e == row).all(1))[0]) for row in np.eye(3, dtype=int)]
    slack = SLACK + (2e-8 if storage == "f32" else 0.0)  # f32: w + d rounds at ~ulp(1/3)/2
    bm = 0.5 * _ulp(storage, d1[0]) + slack
    assert np.all(np.abs(dm) <= bm), (storage, float(np.abs(dm).max()), float(bm.max()))
    for ax in range(3):
        bj = 0.5 * _ulp(storage, d1[qa[ax]]) + slack
        assert np.all(np.abs(dj[ax]) <= bj), (storage, ax, float(np.abs(dj[ax]).max()))
    print(f"{storage}: max |mass miss| {np.abs(dm).max():.2e}, "
          f"max |momentum miss| {np.abs(dj).max():.2e}", flush=True)
