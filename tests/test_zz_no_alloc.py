"""lbm_conservation_static: the FP64 conservation sums with no workspace
argument (a static per-GPU scratch array), through the C ABI, against
lbm_conservation with a caller-allocated workspace. The two share their
kernels, so the results agree bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _both(shape, lattice, storage, seed):
    from maddening_amd import _lib
    from maddening_amd.lbm import LBMKernel, _stream
    k = LBMKernel(shape, lattice=lattice, tau=0.8, storage=storage)
    rng = np.random.default_rng(seed)
    D = len(shape)
    rho = torch.from_numpy(1 + 0.02 * rng.standard_normal(shape)).float()
    u = torch.from_numpy(0.04 * rng.standard_normal(shape + (D,))).float()
    f = k.equilibrium(rho, u)
    d = k.desc()
    L = _lib.lib()
    s = _stream(k.device)
    work = torch.empty(L.lbm_reduce_workspace_size(ctypes.byref(d)), dtype=torch.uint8,
                       device="cuda")
    a = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
    b = torch.full((4,), float("nan"), dtype=torch.float64, device="cuda")
    _lib.check(L.lbm_conservation(s, ctypes.byref(d), _ptr(f), _ptr(work), _ptr(a)),
               "lbm_conservation")
    _lib.check(L.lbm_conservation_static(s, ctypes.byref(d), _ptr(f), _ptr(b)),
               "lbm_conservation_static")
    torch.cuda.synchronize()
    # the state's exact values on the host, as the kernel forms them (float64
    # sums of the stored value, or of the stored deviation plus the float64
    # weight): float64 and f32dev storage only
    exact = None
    if storage == "f64":
        exact = k.to_aos(f, torch.float64).cpu().numpy()
    elif storage == "f32dev":
        w = torch.as_tensor(k.lat.w, dtype=torch.float64).view(-1, *([1] * D))
        exact = (f.cpu().double() + w).permute(*range(1, D + 1), 0).numpy()
    return a.cpu(), b.cpu(), exact


def test_static_scratch_equals_workspace_on_two_calls():
    """A large lattice (every one of the 1024 partial slots used), then a
    smaller one (fewer slots: the sums must not pick up the first call's
    leftovers), then a D2Q9 lattice."""
    cases = [((96, 80, 64), "D3Q19", "f32dev", 1), ((9, 7, 5), "D3Q19", "f64", 2),
             ((33, 17), "D2Q9", "f32", 3)]
    for shape, lattice, storage, seed in cases:
        print(f"static conservation: {shape} {lattice} {storage}", flush=True)
        a, b, f = _both(shape, lattice, storage, seed)
        assert torch.equal(a, b), (shape, a, b)
        if f is None:
            continue
        # and against the sums of the same state on the host
        Q = f.shape[-1]
        from oracle import lbm_numpy as orc
        e, _, _ = orc.lattice(Q)
        mass = f.sum()
        mom = (f.reshape(-1, Q) @ e.astype(np.float64)).sum(axis=0)
        tol = 1e-13 * f.size
        assert abs(float(b[0]) - mass) < tol
        np.testing.assert_allclose(b[1:1 + len(shape)].numpy(), mom, rtol=0, atol=tol)


def test_static_rejects_null_pointers():
    from maddening_amd import _lib
    from maddening_amd.lbm import LBMKernel
    k = LBMKernel((4, 4, 4), tau=0.8)
    d = k.desc()
    L = _lib.lib()
    out = torch.empty(4, dtype=torch.float64, device="cuda")
    assert L.lbm_conservation_static(None, ctypes.byref(d), None, _ptr(out)) == -1
    assert b"non-NULL" in L.lbm_last_error()
