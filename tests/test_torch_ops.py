"""torch.ops.maddening_amd.lbm_step: the LBM step registered as a PyTorch
custom operator. The fake implementation gives shapes without a GPU; on the
GPU the operator equals LBMKernel.step bit for bit, its autograd equals the
adjoint rollout's, and torch.library.opcheck accepts the registration."""
import numpy as np
import pytest
import torch

import maddening_amd.torch_ops  # noqa: F401  (registers the op)

OP = torch.ops.maddening_amd.lbm_step


def test_fake_shapes_without_gpu():
    from torch._subclasses.fake_tensor import FakeTensorMode
    with FakeTensorMode():
        out = OP(torch.empty(19, 4, 5, 6), None, None, None, 0.8, "none", 3, [0.0, 0.0, 0.0])
        assert [tuple(o.shape) for o in out] == [(19, 4, 5, 6), (4, 5, 6), (4, 5, 6, 3), (4, 5, 6)]
        out = OP(torch.empty(9, 7, 8, dtype=torch.float64), None, None, None, 0.8, "none", 1,
                 [0.0, 0.0, 0.0])
        assert out[0].dtype == torch.float64 and out[2].shape == (7, 8, 2)


def _inputs(shape, dtype, seed=0):
    from maddening_amd.lbm import LBMKernel
    rng = np.random.default_rng(seed)
    k = LBMKernel(shape, tau=0.8, storage="f64" if dtype == torch.float64 else "f32",
                  bb="reference")
    u = torch.from_numpy(0.02 * rng.standard_normal(shape + (3,))).float().cuda()
    f = k.equilibrium(torch.ones(shape, device="cuda"), u)
    mask = torch.from_numpy((rng.random(shape) < 0.15).astype(np.uint8)).cuda()
    tau = torch.from_numpy(rng.uniform(0.7, 1.1, shape)).float().cuda()
    return k, f, mask, tau


@pytest.mark.gpu
def test_op_equals_kernel_step():
    k, f, mask, tau = _inputs((10, 9, 12), torch.float32)
    force = torch.tensor([1e-5, 0.0, -2e-5])
    a = OP(f, mask, force, tau, 0.8, "reference", 7, [0.0, 0.0, 0.0])
    b = k.step(f, 7, mask=mask, force=force.tolist(), tau_field=tau)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.gpu
def test_op_gradients_equal_rollout():
    from maddening_amd.adjoint import rollout
    k, f, mask, tau = _inputs((6, 7, 8), torch.float64, seed=1)
    w = torch.randn((6, 7, 8, 3), dtype=torch.float32, device="cuda")

    def loss(outs):
        return (outs[2].double() * w.double()).sum() + outs[0].sum()

    F1 = torch.tensor([1e-5, 2e-5, 0.0], dtype=torch.float64, requires_grad=True)
    t1 = tau.clone().requires_grad_(True)
    f1 = f.clone().requires_grad_(True)
    loss(OP(f1, mask, F1, t1, 0.8, "reference", 5, [0.0, 0.0, 0.0])).backward()
    F2 = F1.detach().cuda().requires_grad_(True)
    t2 = tau.clone().requires_grad_(True)
    f2 = f.clone().requires_grad_(True)
    loss(rollout(k, f2, 5, force=F2, tau_field=t2, mask=mask)).backward()
    torch.testing.assert_close(f1.grad, f2.grad, rtol=1e-12, atol=0)
    torch.testing.assert_close(t1.grad, t2.grad, rtol=1e-6, atol=0)
    torch.testing.assert_close(F1.grad, F2.grad.cpu(), rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_opcheck():
    _, f, mask, tau = _inputs((6, 5, 4), torch.float64, seed=2)
    args = (f, mask, torch.tensor([1e-5, 0.0, 0.0]), tau, 0.8, "reference", 3, [0.0, 0.0, 0.0])
    torch.library.opcheck(OP, args, test_utils=("test_schema", "test_autograd_registration",
                                                "test_faketensor"))


@pytest.mark.gpu
def test_op_backward_does_not_rerun_the_forward(monkeypatch):
    """With an input that requires grad the op's forward is the rollout's
    forward; its kept states reach the backward, which then makes exactly
    the launches of the rollout's own backward (segment recomputes and
    transposed steps), not a second forward pass. The outputs still equal
    LBMKernel.step bit for bit."""
    from maddening_amd import adjoint
    k, f, mask, tau = _inputs((8, 6, 10), torch.float32, seed=3)
    b = k.step(f, 9, mask=mask, force=[1e-5, 0.0, 0.0], tau_field=tau)
    calls = []
    real = adjoint._pull
    monkeypatch.setattr(adjoint, "_pull", lambda *a, **kw: (calls.append(a[3]), real(*a, **kw)))

    def run(op):
        t = tau.clone().requires_grad_(True)
        if op:
            outs = OP(f, mask, torch.tensor([1e-5, 0.0, 0.0]), t, 0.8, "reference", 9,
                      [0.0, 0.0, 0.0])
            assert not hasattr(outs[0], "_lbm_rollout")     # handed to the backward
        else:
            outs = adjoint.rollout(k, f, 9, force=torch.tensor([1e-5, 0.0, 0.0]), tau_field=t,
                                   mask=mask)
        n0 = len(calls)
        outs[2].sum().backward()
        torch.cuda.synchronize()
        return outs, calls[n0:], t.grad
    outs_op, back_op, g_op = run(True)
    outs_ro, back_ro, g_ro = run(False)
    for x, y in zip(outs_op, b):
        assert torch.equal(x, y)
    assert back_op == back_ro
    torch.testing.assert_close(g_op, g_ro, rtol=0, atol=0)
