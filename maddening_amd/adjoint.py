"""Differentiable LBM rollouts: torch.autograd over the HIP adjoint kernels.

``rollout(kern, f0, nsteps, ...)`` advances ``nsteps`` LBMNode.update steps
(lbm.py:899-986) and returns (f_N, rho_N, u_N, p_N). Its backward pass is the
discrete adjoint: the moments' cotangents are pulled back onto f_N
(lbm_moments_adjoint), then each step is transposed (lbm_adjoint_step) from
the last to the first. States the backward needs are recomputed from
checkpoints: with ``checkpoint="sqrt"`` every K = ceil(sqrt(N))-th state is
kept during the forward pass and each K-step segment is recomputed once
(N/K + K states resident, about N extra launches); ``checkpoint="all"``
keeps every state (N states resident) and recomputes nothing; an int sets K;
``checkpoint="auto"`` (the default) keeps every state when N states fit in a
fraction of the device memory that is free, else uses sqrt. Keeping every
state is what jax.grad through the reference's lax.scan does; config 4's
500 steps of 128x128x256 float32 take 159 GB, which fits one MI355X's
288 GB, and then the adjoint costs 2.2x the forward instead of 3.6x.
The forward pass itself is exactly lbm_step's N + 1 launches.

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


# share of the free device memory "auto" lets the N states take
AUTO_MEMORY_SHARE = 0.6


def resolve_checkpoint(state_bytes: int, nsteps: int, checkpoint: Union[str, int],
                       free_bytes: Optional[int] = None, device=None) -> Union[str, int]:
    """``checkpoint`` with "auto" replaced by "all" when nsteps states of
    ``state_bytes`` fit in AUTO_MEMORY_SHARE of ``free_bytes`` (default: the
    device's free memory now), else by "sqrt"."""
    if checkpoint != "auto":
        return checkpoint
    if free_bytes is None:
        free_bytes = torch.cuda.mem_get_info(device)[0]
    return "all" if nsteps * state_bytes <= AUTO_MEMORY_SHARE * free_bytes else "sqrt"


def default_segment(kern: LBMKernel, nsteps: int) -> int:
    """Segment length of the default policy ("auto") for this lattice."""
    state_bytes = kern.ncells * kern.Q * torch.empty((), dtype=kern.dtype).element_size()
    return segment_length(nsteps, resolve_checkpoint(state_bytes, nsteps, "auto",
                                                     device=kern.device))


def segment_length(nsteps: int, checkpoint: Union[str, int]) -> int:
    if checkpoint == "sqrt":
        return max(1, math.ceil(math.sqrt(nsteps)))
    if checkpoint == "all":
        return 1
    k = int(checkpoint)
    if k < 1:
        raise ValueError("checkpoint segment length must be >= 1")
    return k


def _pull(L, st, d, phase, src, dst, m, links, ff, tf, rho=None, u=None, p=None):
    _lib.check(L.lbm_pull_phase(st, ctypes.byref(d), phase, _ptr(src), _ptr(dst), _ptr(m),
                                _ptr(links), _ptr(ff), _ptr(tf), _ptr(rho), _ptr(u), _ptr(p)),
               "lbm_pull_phase")


def _links(L, st, d, kern, m):
    """Halfway link words of the mask (lbm_links_build), read by every
    launch of the rollout in place of the neighbours' mask bytes."""
    if kern.bb != "halfway":
        return None
    links = torch.empty(L.lbm_links_size(ctypes.byref(d)) // 4, dtype=torch.int32,
                        device=kern.device)
    _lib.check(L.lbm_links_build(st, ctypes.byref(d), _ptr(m), _ptr(links)), "lbm_links_build")
    return links


def _two_step(kern: LBMKernel, d: _lib.LbmDesc, fields: bool) -> bool:
    """Whether lbm_step runs this call's sequences (descriptor d: walls,
    force, faces) two units per launch (the same choice, so the rollout's
    forward launches match its own)."""
    if kern.D != 3:
        return False
    nl, npair = ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(_lib.lib().lbm_step_schedule(ctypes.byref(d), 2, int(bool(fields)),
                                            ctypes.byref(nl), ctypes.byref(npair)),
               "lbm_step_schedule")
    return npair.value > 0


def _advance(L, st, d, kern, f0, wanted, nsteps, m, links, ff, tf, pairs, fN, moments):
    """The forward pass of an nsteps rollout that materialises only the
    post-collision states g_k with k in `wanted` (yielded as (k, g) in
    order), then writes f_N and its moments. Units are the pull sequence's:
    g_0 = C(f_0), g_k = C(S B g_{k-1}), f_N = S B g_{N-1}. With `pairs`,
    runs of units between wanted states go two per launch (lbm_pull_phase 3
    (SC, SC), 4 (C, SC), 5 (SC, S), 6 (C, S)), bit-identical to one per
    launch."""
    pool = [kern.empty(), kern.empty()]
    cur, idx = f0, -1          # f_0 (idx -1) or g_idx
    nslot = 0

    def scratch():
        nonlocal nslot
        nslot += 1
        return pool[nslot & 1]

    stops = list(wanted)
    last = nsteps - 2 if pairs else nsteps - 1   # the final launch starts from g_last
    if last >= 0 and (not stops or stops[-1] != last):
        stops.append(last)
    for t in stops:
        while idx < t:
            two = pairs and t - idx >= 2
            dst = kern.empty() if (idx + (2 if two else 1)) == t and t in wanted else scratch()
            if two:
                _pull(L, st, d, 4 if idx < 0 else 3, cur, dst, m, links, ff, tf)
                idx += 2
            else:
                _pull(L, st, d, 0 if idx < 0 else 1, cur, dst, m, links, ff, tf)
                idx += 1
            cur = dst
        if t in wanted:
            yield t, cur
    rho, u, p = moments
    if pairs:
        _pull(L, st, d, 6 if idx < 0 else 5, cur, fN, m, links, ff, tf, rho, u, p)
    else:
        _pull(L, st, d, 2, cur, fN, m, links, ff, tf, rho, u, p)


def _forward_impl(ctx, f0, force, tau_field, rho_in, rho_out, kern: LBMKernel, nsteps: int,
                  mask, seg: int, face_names):
    """The rollout's forward pass: runs the pull sequence, keeps the states
    the backward needs (checkpoints and the last segment), and records what
    the backward needs on `ctx` (an autograd context or any object).
    Returns ((f_N, rho, u, p), tensors to keep for the backward)."""
    L = _lib.lib()
    dev = kern.device
    st = _stream(dev)
    # a uniform force goes to the kernels as host numbers: read it once
    fk = force.tolist() if (force is not None and force.dim() == 1) else force
    ff, fu = kern._field(fk)
    faces = _faces(face_names, rho_in, rho_out)
    d = kern.desc(force=fu, faces=faces)
    m = kern._mask(mask)
    if kern.bb != "none" and m is None:
        raise ValueError(f"bb='{kern.bb}' needs a wall mask")
    tf = None if tau_field is None else tau_field.to(dev, torch.float32).contiguous()
    links = _links(L, st, d, kern, m)
    ckpts = [f0]           # source of each segment's first state
    # the last segment's states come straight from this pass: no recompute
    s_last = ((nsteps - 1) // seg) * seg
    keep = {k for k in range(nsteps - 1) if (k + 1) % seg == 0}   # g_k starts segment k + 1
    tail_idx = set(range(s_last, nsteps - 1))
    fN = kern.empty()
    rho = torch.empty(kern.grid_shape, dtype=torch.float32, device=dev)
    u = torch.empty(kern.grid_shape + (kern.D,), dtype=torch.float32, device=dev)
    p = torch.empty(kern.grid_shape, dtype=torch.float32, device=dev)
    pairs = _two_step(kern, d, ff is not None or tf is not None)
    tail = []
    for k, g in _advance(L, st, d, kern, f0, sorted(keep | tail_idx), nsteps, m, links, ff,
                         tf, pairs, fN, (rho, u, p)):
        if k in keep:
            ckpts.append(g)
        if k in tail_idx:
            tail.append(g)
    ctx.kern = kern
    ctx.links = links
    ctx.nsteps = nsteps
    ctx.seg = seg
    ctx.mask = mask
    ctx.force_is_field = force is not None and force.dim() > 1
    ctx.force_dtype = None if force is None else force.dtype
    ctx.force_list = fk if isinstance(fk, list) else None
    ctx.n_ckpt = len(ckpts)
    ctx.faces = faces
    ctx.rho_dtypes = (None if rho_in is None else rho_in.dtype,
                      None if rho_out is None else rho_out.dtype)
    return (fN, rho, u, p), [fN, force, tau_field, *ckpts, *tail]


def _backward_impl(ctx, saved, needs, g_f, g_rho, g_u, g_p):
    """The rollout's backward pass from what _forward_impl recorded on `ctx`
    and kept in `saved`; `needs` flags the gradients wanted, for (f_0,
    force, tau_field, rho_in, rho_out). Returns those five gradients (None
    where not wanted) followed by five Nones."""
    kern: LBMKernel = ctx.kern
    f_N, force, tau_field, *rest = saved
    ckpts, tail = rest[:ctx.n_ckpt], rest[ctx.n_ckpt:]
    L = _lib.lib()
    dev = kern.device
    st = _stream(dev)
    if ctx.force_list is not None:
        ff, fu = None, ctx.force_list
    else:
        ff, fu = kern._field(force)
    d = kern.desc(force=fu, faces=ctx.faces)
    m = kern._mask(ctx.mask)
    tf = None if tau_field is None else tau_field.to(dev, torch.float32).contiguous()
    need_f0, need_force, need_tau = needs[0], needs[1], needs[2]
    work = torch.empty(L.lbm_adjoint_workspace_size(ctypes.byref(d)), dtype=torch.uint8,
                       device=dev)
    _lib.check(L.lbm_adjoint_begin(st, ctypes.byref(d), _ptr(work)), "lbm_adjoint_begin")
    tau_bar = torch.zeros(kern.grid_shape, dtype=torch.float32, device=dev) \
        if (need_tau and tf is not None) else None
    fbar_field = torch.zeros(kern.grid_shape + (kern.D,), dtype=torch.float32, device=dev) \
        if (need_force and ctx.force_is_field) else None

    def c(t):
        return None if t is None else t.to(dev, torch.float32).contiguous()

    # Under the reference wall rule the cotangents between steps are kept
    # with the slots of wall cells exchanged, so each transposed step is a
    # plain gather (no neighbour mask reads); the last step writes dL/df_0
    # in the ordinary form.
    # With Zou-He faces every cotangent but dL/df_0 is passed back through
    # the faces' transpose (their states came out of the faces), which
    # also accumulates dL/drho of each face.
    faces_on = ctx.faces is not None
    OUTF = _lib.LBM_FLAG_ADJ_OUT_FACES if faces_on else 0

    def with_flags(bits, extra=0):
        dd = _lib.LbmDesc.from_buffer_copy(d)
        if kern.bb == "reference":
            dd.flags |= bits
        dd.flags |= extra
        return dd
    d_mom = with_flags(_lib.LBM_FLAG_ADJ_OUT_SWAPPED, OUTF)
    d_mid = with_flags(_lib.LBM_FLAG_ADJ_IN_SWAPPED | _lib.LBM_FLAG_ADJ_OUT_SWAPPED, OUTF)
    d_first = with_flags(_lib.LBM_FLAG_ADJ_IN_SWAPPED)
    fbar = torch.empty(kern.state_shape, dtype=kern.dtype, device=dev)
    gin = None if g_f is None else g_f.to(dev, kern.dtype).contiguous()
    _lib.check(L.lbm_moments_adjoint(st, ctypes.byref(d_mom), _ptr(f_N), _ptr(m), _ptr(ff),
                                     _ptr(c(g_rho)), _ptr(c(g_u)), _ptr(c(g_p)), _ptr(gin),
                                     _ptr(fbar), _ptr(fbar_field), _ptr(work)),
               "lbm_moments_adjoint")
    # Per segment (last first): recompute g_s .. g_{s+n-2} from the
    # checkpoint with the forward's pull launches, then transpose steps
    # s+n-1 .. s; step k's starting state is f_0 or gathered from g_{k-1}.
    spare = torch.empty_like(fbar)
    for si in reversed(range(len(ckpts))):
        s = si * ctx.seg
        n = min(ctx.seg, ctx.nsteps - s)
        states = [ckpts[si]]
        if si == len(ckpts) - 1:
            states += list(tail)
        for i in range(len(states), n):
            g = kern.empty()
            _pull(L, st, d, 1 if (s > 0 or i > 1) else 0, states[-1], g, m, ctx.links, ff, tf)
            states.append(g)
        for i in reversed(range(n)):
            post = int(s > 0 or i > 0)
            dk = d_mid if post else d_first
            _lib.check(L.lbm_adjoint_step(st, ctypes.byref(dk), _ptr(states[i]), post,
                                          _ptr(fbar), _ptr(spare), _ptr(m),
                                          _ptr(ctx.links), _ptr(ff), _ptr(tf),
                                          _ptr(tau_bar), _ptr(fbar_field), _ptr(work)),
                       "lbm_adjoint_step")
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
    grad_rin = grad_rout = None
    if faces_on and (needs[3] or needs[4]):
        rb = torch.empty(2, dtype=torch.float64, device=dev)
        _lib.check(L.lbm_adjoint_face_end(st, ctypes.byref(d), _ptr(work), _ptr(rb)),
                   "lbm_adjoint_face_end")
        if needs[3]:
            grad_rin = rb[0].to(ctx.rho_dtypes[0])
        if needs[4]:
            grad_rout = rb[1].to(ctx.rho_dtypes[1])
    return ((fbar if need_f0 else None), grad_force, grad_tau, grad_rin, grad_rout,
            None, None, None, None, None)


class _Rollout(torch.autograd.Function):
    """The forward pass runs the pull sequence of lbm_step launch by launch:
    g_0 = C(f_0), g_k = C(S B g_{k-1}), f_N = S B g_{N-1}. The starting state
    of step k is f_0 (k = 0) or is gathered from g_{k-1}. A segment starting
    at step s is therefore checkpointed by keeping f_0 (s = 0) or g_{s-1}."""

    @staticmethod
    def forward(ctx, f0, force, tau_field, rho_in, rho_out, kern: LBMKernel, nsteps: int, mask,
                seg: int, face_names):
        ctx.set_materialize_grads(False)
        outs, keep = _forward_impl(ctx, f0, force, tau_field, rho_in, rho_out, kern, nsteps, mask,
                                   seg, face_names)
        ctx.save_for_backward(*keep)
        return outs

    @staticmethod
    def backward(ctx, g_f, g_rho, g_u, g_p):
        return _backward_impl(ctx, ctx.saved_tensors, ctx.needs_input_grad, g_f, g_rho, g_u, g_p)


def _faces(names, rho_in, rho_out):
    """((inlet, rho_in), (outlet, rho_out)) for LBMKernel.desc, or None."""
    if names is None:
        return None
    vals = [None if r is None else float(r) for r in (rho_in, rho_out)]
    out = tuple(None if n is None else (n, v) for n, v in zip(names, vals))
    return None if out == (None, None) else out


def rollout(kern: LBMKernel, f0: torch.Tensor, nsteps: int, *,
            force: Optional[torch.Tensor] = None, tau_field: Optional[torch.Tensor] = None,
            mask: Optional[torch.Tensor] = None, checkpoint: Union[str, int] = "auto",
            faces=None):
    """Differentiable ``nsteps``-step rollout; returns (f_N, rho, u, p).

    ``force`` is None, a (D,) tensor (uniform) or a (*G, D) tensor;
    ``tau_field`` is None or a (*G) tensor; either may require grad, as may
    ``f0``. ``faces``: Zou-He pressure faces ((inlet_face, rho_in),
    (outlet_face, rho_out)), either entry None, with each rho (= pressure /
    cs2) a float or a 0-dim tensor that may require grad. Needs f32, f32dev
    or f64 storage.
    """
    if kern.storage not in ("f32", "f32dev", "f64"):
        raise ValueError("differentiable rollouts need f32, f32dev or f64 storage")
    if nsteps < 1:
        raise ValueError("nsteps must be >= 1")
    if force is not None and not torch.is_tensor(force):
        force = torch.as_tensor(force, dtype=torch.float64, device=kern.device)
    state_bytes = kern.ncells * kern.Q * torch.empty((), dtype=kern.dtype).element_size()
    checkpoint = resolve_checkpoint(state_bytes, nsteps, checkpoint, device=kern.device)
    seg = segment_length(nsteps, checkpoint)
    names, rho_in, rho_out = None, None, None
    if faces is not None:
        names = tuple(None if fc is None else fc[0] for fc in faces)

        def as_t(fc):
            if fc is None:
                return None
            v = fc[1]
            return v if torch.is_tensor(v) else torch.tensor(float(v), dtype=torch.float64)
        rho_in, rho_out = as_t(faces[0]), as_t(faces[1])
    return _Rollout.apply(f0, force, tau_field, rho_in, rho_out, kern, int(nsteps), mask, seg,
                          names)
