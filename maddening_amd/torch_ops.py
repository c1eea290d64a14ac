"""The LBM step as a registered PyTorch operator: ``torch.ops.maddening_amd.lbm_step``.

This is the PyTorch counterpart of the jax.ffi target the reference's
``LBMNode.update`` would bind (lbm.py:899-986). It is a
``torch.library.custom_op`` (functional: fresh outputs, no aliasing) with:

* a fake implementation, so shape propagation and ``torch.compile``
  tracing see the op without running it;
* an autograd formula. It returns the gradients of f_0, of the force
  (uniform (D,) or per cell (*G, D)) and of the per-cell tau field through
  the HIP adjoint (``maddening_amd.adjoint.rollout``, sqrt(N)
  checkpointing), recomputing the forward from the saved inputs.

Arguments: ``f`` is the SoA state (Q, *G), float32 or float64. ``mask``
is an optional (*G) uint8 wall mask. ``force`` is optional, (D,) or
(*G, D). ``tau_field`` is optional, (*G). ``tau`` is the scalar
relaxation time used where no field is given. ``bb`` is "none",
"reference" or "halfway". ``lid_velocity`` has 3 entries. Returns
(f_N, rho, u, p); the moments are float32.
"""
from __future__ import annotations

from typing import Optional

import torch

from .lbm import LBMKernel

_KERNELS: dict = {}


def _kernel(f: torch.Tensor, tau: float, bb: str, lid: list) -> LBMKernel:
    q = int(f.shape[0])
    key = (tuple(f.shape), f.dtype, f.device, float(tau), bb, tuple(float(v) for v in lid))
    k = _KERNELS.get(key)
    if k is None:
        storage = "f64" if f.dtype == torch.float64 else "f32"
        k = LBMKernel(tuple(int(n) for n in f.shape[1:]), lattice="D3Q19" if q == 19 else "D2Q9",
                      tau=tau, storage=storage, bb=bb, lid_velocity=tuple(lid), device=f.device)
        _KERNELS[key] = k
    return k


def _force_arg(force: Optional[torch.Tensor]):
    if force is None:
        return None
    return [float(v) for v in force.tolist()] if force.dim() == 1 else force


@torch.library.custom_op("maddening_amd::lbm_step", mutates_args=())
def lbm_step(f: torch.Tensor, mask: Optional[torch.Tensor], force: Optional[torch.Tensor],
             tau_field: Optional[torch.Tensor], tau: float, bb: str, nsteps: int,
             lid_velocity: list[float]) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor,
                                                 torch.Tensor]:
    k = _kernel(f, tau, bb, lid_velocity)
    fN, rho, u, p = k.step(f.contiguous(), int(nsteps), mask=mask, force=_force_arg(force),
                           tau_field=tau_field)
    return fN, rho, u, p


@lbm_step.register_fake
def _(f, mask, force, tau_field, tau, bb, nsteps, lid_velocity):
    G = tuple(f.shape[1:])
    D = len(G)
    return (torch.empty_like(f), f.new_empty(G, dtype=torch.float32),
            f.new_empty(G + (D,), dtype=torch.float32), f.new_empty(G, dtype=torch.float32))


def _setup_context(ctx, inputs, output):
    f, mask, force, tau_field, tau, bb, nsteps, lid = inputs
    ctx.save_for_backward(f, mask, force, tau_field)
    ctx.args = (float(tau), bb, int(nsteps), list(lid))


def _backward(ctx, g_f, g_rho, g_u, g_p):
    from .adjoint import rollout
    f, mask, force, tau_field = ctx.saved_tensors
    tau, bb, nsteps, lid = ctx.args
    k = _kernel(f, tau, bb, lid)
    need_f, _, need_force, need_tau = ctx.needs_input_grad[:4]
    with torch.enable_grad():
        f0 = f.detach().requires_grad_(need_f)
        fr = None
        if force is not None:
            fr = force.detach().to(torch.float64 if force.dim() == 1 else force.dtype)
            fr = fr.to(k.device).requires_grad_(need_force)
        tf = None if tau_field is None else tau_field.detach().requires_grad_(need_tau)
        outs = rollout(k, f0, nsteps, force=fr, tau_field=tf, mask=mask)
        pairs = [(o, g) for o, g in zip(outs, (g_f, g_rho, g_u, g_p)) if g is not None]
        wrt = [t for t in (f0, fr, tf) if t is not None and t.requires_grad]
        grads = torch.autograd.grad([o for o, _ in pairs], wrt, [g for _, g in pairs],
                                    allow_unused=True) if pairs and wrt else []
    it = iter(grads)
    gf = next(it) if f0.requires_grad else None
    gF = next(it) if (fr is not None and fr.requires_grad) else None
    gT = next(it) if (tf is not None and tf.requires_grad) else None
    if gF is not None and force is not None:
        gF = gF.to(device=force.device, dtype=force.dtype)
    return gf, None, gF, gT, None, None, None, None


lbm_step.register_autograd(_backward, setup_context=_setup_context)
