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


@pytest.mark.parametrize("shape,faces", [((12, 9), (("x_min", 1.03), ("x_max", 0.99))),
                                         ((6, 7, 8), (("z_min", 1.02), ("y_max", 0.98)))])
def test_torch_zou_he_faces_match_numpy_and_rho_gradient(shape, faces):
    """Zou-He faces in the torch restatement: values equal the NumPy oracle's
    (which takes pressure = rho cs2), and d/d rho_p of a loss equals central
    differences."""
    from oracle import lbm_numpy as on
    rng = np.random.default_rng(11)
    Q = 19 if len(shape) == 3 else 9
    e, w, _ = on.lattice(Q)
    D = len(shape)
    f0 = on.equilibrium(1 + 0.01 * rng.random(shape), 0.02 * rng.standard_normal(shape + (D,)),
                        e, w)
    mask = (rng.random(shape) < 0.15).astype(np.uint8)
    a = on.update(f0, tau=0.8, mask=mask, inlet=(faces[0][0], faces[0][1] / 3),
                  outlet=(faces[1][0], faces[1][1] / 3))
    zero = torch.zeros(D, dtype=torch.float64)
    b, _, _ = ot.update(torch.from_numpy(f0), tau=0.8, force=zero, mask=torch.from_numpy(mask),
                        inlet=faces[0], outlet=faces[1])
    assert np.abs(a["f"] - b.numpy()).max() < 1e-15
    wts = torch.from_numpy(rng.standard_normal(shape + (Q,)))

    def loss(r_in, r_out):
        out, rho, u = ot.update(torch.from_numpy(f0), tau=0.8, force=zero,
                                mask=torch.from_numpy(mask), inlet=(faces[0][0], r_in),
                                outlet=(faces[1][0], r_out))
        return (out * wts).sum() + (u * wts[..., :D]).sum()

    r = [torch.tensor(faces[0][1], dtype=torch.float64, requires_grad=True),
         torch.tensor(faces[1][1], dtype=torch.float64, requires_grad=True)]
    loss(*r).backward()
    h = 1e-6
    for i in range(2):
        plus = [float(x.detach()) for x in r]
        minus = [float(x.detach()) for x in r]
        plus[i] += h
        minus[i] -= h
        fd = (float(loss(*plus)) - float(loss(*minus))) / (2 * h)
        assert abs(float(r[i].grad) - fd) < 1e-6 * max(1.0, abs(fd)), (i, float(r[i].grad), fd)
