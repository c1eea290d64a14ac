"""Torch float64 restatement of the reference LBM step, for gradients: ORACLE.

Test infrastructure only (see oracle/lbm_numpy.py). It stands in for
``jax.grad`` of the reference (JAX is not available here): the same
operations as ``lbm_numpy.update`` (lbm.py:899-986), written with torch ops
so that autograd differentiates them. tests/test_oracle.py checks it against
``lbm_numpy`` value for value; its gradients are checked against central
finite differences there too.
"""
from __future__ import annotations

import torch

from . import lbm_numpy as _np_oracle

CS2 = 1.0 / 3.0


def lattice(Q, device="cpu"):
    e, w, opp = _np_oracle.lattice(Q)
    return (torch.as_tensor(e, dtype=torch.float64, device=device),
            torch.as_tensor(w, dtype=torch.float64, device=device),
            torch.as_tensor(opp, dtype=torch.long, device=device), e)


def equilibrium(rho, u, e, w):
    """lbm.py:180-206."""
    eu = u @ e.T
    usq = (u * u).sum(-1, keepdim=True)
    return w * rho[..., None] * (1 + eu / CS2 + eu ** 2 / (2 * CS2 ** 2) - usq / (2 * CS2))


def macroscopic(f, e, force):
    """lbm.py:209-239."""
    rho = f.sum(-1)
    return rho, (f @ e + 0.5 * force) / torch.clamp(rho[..., None], min=1e-10)


def guo_forcing(u, force, tau, e, w):
    """lbm.py:307-353 (tau scalar or per-cell)."""
    eu = u @ e.T
    bracket = (e[None, :, :] - u[..., None, :]) / CS2 + e * (eu / CS2 ** 2)[..., None]
    pref = 1 - 0.5 / tau
    if torch.is_tensor(pref) and pref.dim():
        pref = pref[..., None]
    return pref * w * (bracket * force[..., None, :]).sum(-1)


def stream(f, e_int):
    """lbm.py:242-267: f_q(x) <- f_q(x - e_q), periodic."""
    outs = []
    for q in range(e_int.shape[0]):
        fq = f[..., q]
        shifts = [int(s) for s in e_int[q]]
        dims = [d for d, s in enumerate(shifts) if s]
        if dims:
            fq = torch.roll(fq, [shifts[d] for d in dims], dims)
        outs.append(fq)
    return torch.stack(outs, -1)


def zou_he_pressure_face(f, rho_p, e, w, opp, face, wall):
    """Non-equilibrium bounce-back on one face (lbm.py:425-510), with the
    reference's normal velocity u_n = +-(1 - (sum known + sum tang)/rho_p);
    rho_p may be a float or a 0-dim tensor (then it is differentiable)."""
    axis, side = _np_oracle.FACES[face]
    D = e.shape[1]
    known, unknown, tang = _np_oracle.classify_directions(e.cpu().numpy().astype(int), axis, side)
    idx = [slice(None)] * D
    idx[axis] = 0 if side == "min" else -1
    idx = tuple(idx)
    ff = f[idx]
    wf = wall[idx]
    s = ff[..., known].sum(-1) + ff[..., tang].sum(-1)
    rho_p = torch.as_tensor(rho_p, dtype=f.dtype, device=f.device)
    r = torch.clamp(rho_p, min=1e-10)
    un = 1.0 - s / r if side == "min" else -1.0 + s / r
    vel = torch.zeros(ff.shape[:-1] + (D,), dtype=f.dtype, device=f.device)
    vel = torch.cat([vel[..., :axis], un[..., None], vel[..., axis + 1:]], -1)
    feq = equilibrium(rho_p.expand(ff.shape[:-1]), vel, e, w)
    cols = []
    for q in range(f.shape[-1]):
        if q in unknown:
            o = int(opp[q])
            cand = ff[..., o] + feq[..., q] - feq[..., o]
            cols.append(torch.where(wf, ff[..., q], cand))
        else:
            cols.append(ff[..., q])
    new = torch.stack(cols, -1)
    out = f.clone()
    out[idx] = new
    return out


def update(f, *, tau, force, mask=None, bb="reference", lid_velocity=None, inlet=None,
           outlet=None):
    """One step on an AoS (*G, Q) float64 tensor; returns (f, rho, u).
    inlet/outlet: optional (face, rho_p) Zou-He faces (lbm.py:950-968), with
    rho_p = pressure / cs2 a float or a 0-dim tensor."""
    Q = f.shape[-1]
    e, w, opp, e_int = lattice(Q, f.device)
    G = f.shape[:-1]
    D = len(G)
    force = force.expand(G + (D,)) if force.dim() == 1 else force
    rho, u = macroscopic(f, e, force)
    feq = equilibrium(rho, u, e, w)
    t = tau[..., None] if (torch.is_tensor(tau) and tau.dim()) else tau
    f_post = f - (f - feq) / t + guo_forcing(u, force, tau, e, w)
    f_str = stream(f_post, e_int)
    wall = torch.zeros(G, dtype=torch.bool, device=f.device) if mask is None else mask != 0
    if bb == "halfway":
        out = f_str
        lid = torch.zeros(D, dtype=torch.float64) if lid_velocity is None else \
            torch.as_tensor(lid_velocity[:D], dtype=torch.float64)
        cols = [out[..., 0]]
        for q in range(1, Q):
            src = mask
            shifts = [int(s) for s in e_int[q]]
            dims = [d for d, s in enumerate(shifts) if s]
            src = torch.roll(src, [shifts[d] for d in dims], dims)
            hit = (src != 0) & ~wall
            refl = f_post[..., int(opp[q])] + torch.where(
                src == 2, 2 * w[q] * float(e[q] @ lid) / CS2, 0.0)
            cols.append(torch.where(hit, refl, out[..., q]))
        out = torch.stack(cols, -1)
        out = torch.where(wall[..., None], w.expand(out.shape), out)
    else:
        out = torch.where(wall[..., None], f_str[..., opp], f_str)
    for fc in (inlet, outlet):
        if fc is not None:
            out = zou_he_pressure_face(out, fc[1], e, w, opp, fc[0], wall)
    rho_n, u_n = macroscopic(out, e, force)
    u_n = torch.where(wall[..., None], 0.0, u_n)
    return out, rho_n, u_n
