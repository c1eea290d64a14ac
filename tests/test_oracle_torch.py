"""Pins of the torch float64 restatement (oracle/lbm_torch.py) used for
gradients: value-for-value against the NumPy oracle, and its autograd
against central finite differences."""
import numpy as np
import pytest
import torch

from oracle import lbm_numpy as on
from oracle import lbm_torch as ot


@pytest.mark.parametrize("shape", [(4, 5, 6), (6, 7)])
@pytest.mark.parametrize("bb", ["reference", "halfway"])
def test_torch_restatement_matches_numpy_oracle(shape, bb):
    rng = np.random.default_rng(0)
    D = len(shape)
    Q = 19 if D == 3 else 9
    e, w, _ = on.lattice(Q)
    f = on.equilibrium(1 + 0.02 * rng.standard_normal(shape), 0.03 * rng.standard_normal(shape + (D,)), e, w)
    mask = (rng.random(shape) < 0.3).astype(np.uint8)
    mask[mask == 1] = rng.integers(1, 3, int(mask.sum())).astype(np.uint8)
    tau = rng.uniform(0.6, 1.2, shape)
    F = 1e-4 * rng.standard_normal(shape + (D,))
    lid = (0.05, 0.01, -0.02)[:D]
    a = on.update(f, tau=tau, force=F, mask=mask, bb=bb, lid_velocity=lid)
    fb, rb, ub = ot.update(torch.from_numpy(f), tau=torch.from_numpy(tau), force=torch.from_numpy(F),
                           mask=torch.from_numpy(mask), bb=bb, lid_velocity=lid)
    np.testing.assert_allclose(fb.numpy(), a["f"], rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(rb.numpy(), a["density"], rtol=1e-14)
    np.testing.assert_allclose(ub.numpy(), a["velocity"], rtol=1e-12, atol=1e-16)


def test_torch_restatement_gradient_vs_central_differences():
    rng = np.random.default_rng(1)
    shape = (4, 4, 5)
    e, w, _ = on.lattice(19)
    f0 = torch.from_numpy(on.equilibrium(1 + 0.02 * rng.standard_normal(shape),
                                         0.03 * rng.standard_normal(shape + (3,)), e, w))
    mask = torch.from_numpy((rng.random(shape) < 0.2).astype(np.uint8))
    tau = torch.from_numpy(rng.uniform(0.6, 1.2, shape))
    F = torch.tensor([1e-4, -2e-4, 3e-4], dtype=torch.float64)
    wt = torch.from_numpy(rng.standard_normal(shape + (3,)))

    def loss(tau_, F_):
        f = f0
        for _ in range(3):
            f, r, u = ot.update(f, tau=tau_, force=F_, mask=mask)
        return (u * wt).sum()

    t = tau.clone().requires_grad_(True)
    Fg = F.clone().requires_grad_(True)
    loss(t, Fg).backward()
    for d in range(3):
        h = 1e-6
        Fp, Fm = F.clone(), F.clone()
        Fp[d] += h
        Fm[d] -= h
        fd = (loss(tau, Fp) - loss(tau, Fm)) / (2 * h)
        assert float(Fg.grad[d]) == pytest.approx(float(fd), rel=1e-6)
    for c in [(0, 0, 0), (1, 2, 3), (3, 3, 4)]:
        h = 1e-6
        tp, tm = tau.clone(), tau.clone()
        tp[c] += h
        tm[c] -= h
        fd = (loss(tp, F) - loss(tm, F)) / (2 * h)
        assert float(t.grad[c]) == pytest.approx(float(fd), rel=1e-5, abs=1e-9)
