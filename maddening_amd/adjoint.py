"""Differentiable LBM rollouts: torch.autograd over the HIP adjoint kernels.

``rollout(kern, f0, nsteps, ...)`` advances ``nsteps`` LBMNode.update steps
(lbm.py:899-986) and returns (f_N, rho_N, u_N, p_N). Its backward pass is the
discrete adjoint: the moments' cotangents are pulled back onto f_N
(lbm_moments_adjoint), then each step is transposed (lbm_adjoint_step) from
the last to the first. States the backward needs are recomputed from
checkpoints: with ``checkpoint="sqrt"`` every K = ceil(sqrt(N))-th state is
kept during the forward pass and each K-step segment is recomputed once
(N/K + K states resident); ``checkpoint="all"`` keeps every state and
recomputes nothing; an int sets K.

Gradients flow to f0, to a uniform body force (tensor of shape (D,)) or a
force field (*G, D), and to a per-cell relaxation-time field (*G).
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Union

import torch

from . import _lib
from .lbm import LBMKernel, _ptr, _stream


def segment_length(nsteps: int, checkpoint: Union[str, int]) -> int:
    if checkpoint == "sqrt":
        return max(1, math.ceil(math.sqrt(nsteps)))
    if checkpoint == "all":
        return 1
    k = int(checkpoint)
    if k < 1:
        raise ValueError("checkpoint segment length must be >= 1")
    return k


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, force, tau_field, kern: LBMKernel, nsteps: int, mask, seg: int):
        ctx.set_materialize_grads(False)
        # a uniform force goes to the kernels as host numbers: read it once
        fk = force.tolist() if (force is not None and force.dim() == 1) else force
        ckpts = []
        f = f0
        done = 0
        while done < nsteps:
            n = min(seg, nsteps - done)
            ckpts.append(f)
            f, _, _, _ = kern.step(f, n, mask=mask, force=fk, tau_field=tau_field,
                                   moments=False)
            done += n
        rho, u, p = kern.moments(f, mask=mask, force=fk)
        ctx.kern = kern
        ctx.nsteps = nsteps
        ctx.seg = seg
        ctx.mask = mask
        ctx.force_is_field = force is not None and force.dim() > 1
        ctx.force_shape = None if force is None else tuple(force.shape)
        ctx.force_dtype = None if force is None else force.dtype
        ctx.force_list = fk if isinstance(fk, list) else None
        ctx.save_for_backward(f, force, tau_field, *ckpts)
        return f, rho, u, p

    @staticmethod
    def backward(ctx, g_f, g_rho, g_u, g_p):
        kern: LBMKernel = ctx.kern
        f_N, force, tau_field, *ckpts = ctx.saved_tensors
        L = _lib.lib()
        dev = kern.device
        st = _stream(dev)
        if ctx.force_list is not None:
            ff, fu = None, ctx.force_list
        else:
            ff, fu = kern._field(force)
        d = kern.desc(force=fu)
        m = kern._mask(ctx.mask)
        if kern.bb != "none" and m is None:
            raise ValueError("bb mode needs a wall mask")
        tf = None if tau_field is None else tau_field.to(dev, torch.float32).contiguous()
        need_f0, need_force, need_tau = ctx.needs_input_grad[0], ctx.needs_input_grad[1], \
            ctx.needs_input_grad[2]
        work = torch.empty(L.lbm_adjoint_workspace_size(ctypes.byref(d)), dtype=torch.uint8,
                           device=dev)
        _lib.check(L.lbm_adjoint_begin(st, ctypes.byref(d), _ptr(work)), "lbm_adjoint_begin")
        tau_bar = torch.zeros(kern.grid_shape, dtype=torch.float32, device=dev) \
            if (need_tau and tf is not None) else None
        fbar_field = torch.zeros(kern.grid_shape + (kern.D,), dtype=torch.float32, device=dev) \
            if (need_force and ctx.force_is_field) else None

        def c(t):
            return None if t is None else t.to(dev, torch.float32).contiguous()

        fbar = torch.empty(kern.state_shape, dtype=kern.dtype, device=dev)
        gin = None if g_f is None else g_f.to(dev, kern.dtype).contiguous()
        _lib.check(L.lbm_moments_adjoint(st, ctypes.byref(d), _ptr(f_N), _ptr(m), _ptr(ff),
                                         _ptr(c(g_rho)), _ptr(c(g_u)), _ptr(c(g_p)), _ptr(gin),
                                         _ptr(fbar), _ptr(fbar_field), _ptr(work)),
                   "lbm_moments_adjoint")
        # Recompute each segment with the forward's pull launches, keeping the
        # post-collision states g_0 .. g_{n-2}; step k's starting state is the
        # checkpoint for k = 0 and is gathered from g_{k-1} by the adjoint
        # kernel otherwise.
        spare = torch.empty_like(fbar)
        nseg = len(ckpts)
        for s in reversed(range(nseg)):
            n = min(ctx.seg, ctx.nsteps - s * ctx.seg)
            states = [ckpts[s]]
            for k in range(n - 1):
                g = torch.empty(kern.state_shape, dtype=kern.dtype, device=dev)
                _lib.check(L.lbm_pull_phase(st, ctypes.byref(d), 0 if k == 0 else 1,
                                            _ptr(states[-1]), _ptr(g), _ptr(m), _ptr(ff),
                                            _ptr(tf), None, None, None), "lbm_pull_phase")
                states.append(g)
            for i in reversed(range(n)):
                _lib.check(L.lbm_adjoint_step(st, ctypes.byref(d), _ptr(states[i]), int(i > 0),
                                              _ptr(fbar), _ptr(spare), _ptr(m), _ptr(ff),
                                              _ptr(tf), _ptr(tau_bar), _ptr(fbar_field),
                                              _ptr(work)), "lbm_adjoint_step")
                fbar, spare = spare, fbar
            del states
        grad_force = None
        if need_force and force is not None:
            if ctx.force_is_field:
                grad_force = fbar_field.to(ctx.force_dtype)
            else:
                out = torch.empty(3, dtype=torch.float64, device=dev)
                _lib.check(L.lbm_adjoint_end(st, ctypes.byref(d), _ptr(work), _ptr(out)),
                           "lbm_adjoint_end")
                grad_force = out[: kern.D].to(ctx.force_dtype)
        grad_tau = tau_bar.to(tau_field.dtype) if tau_bar is not None else None
        return (fbar if need_f0 else None), grad_force, grad_tau, None, None, None, None


def rollout(kern: LBMKernel, f0: torch.Tensor, nsteps: int, *,
            force: Optional[torch.Tensor] = None, tau_field: Optional[torch.Tensor] = None,
            mask: Optional[torch.Tensor] = None, checkpoint: Union[str, int] = "sqrt"):
    """Differentiable ``nsteps``-step rollout; returns (f_N, rho, u, p).

    ``force`` is None, a (D,) tensor (uniform) or a (*G, D) tensor;
    ``tau_field`` is None or a (*G) tensor; either may require grad, as may
    ``f0``. Needs f32 or f64 storage.
    """
    if kern.storage not in ("f32", "f32dev", "f64"):
        raise ValueError("differentiable rollouts need f32, f32dev or f64 storage")
    if nsteps < 1:
        raise ValueError("nsteps must be >= 1")
    if force is not None and not torch.is_tensor(force):
        force = torch.as_tensor(force, dtype=torch.float64, device=kern.device)
    seg = segment_length(nsteps, checkpoint)
    return _Rollout.apply(f0, force, tau_field, kern, int(nsteps), mask, seg)
