// LBMPipeNode's step on the GPU: D3Q19 fluid in a pipe with an actuator-disc
// propeller and gravity, plus either a D3Q7 passive tracer (single phase) or
// Shan-Chen multiphase forcing (src/maddening/nodes/lbm_pipe.py:751-879).
//
// State layout: SoA f[19][nx][ny][nz] and g[7][nx][ny][nz] (the tracer),
// storing f itself (not deviations) in float or double. Per-cell flags:
// bit 0 wall, bit 1 propeller disc. The force on a cell is
// (propeller ? strength : 0, 0, wall ? 0 : gravity) (lbm_pipe.py:760-769).
//
// Single phase, n steps, pull scheme with post-collision buffers F and G:
//   launch 0        F_0 = C(f_0)
//   launch k (1..n) f_k = B S F_{k-1} is gathered; its moments are step k-1's
//                   outputs (u = 0 at walls). Step k-1's tracer collision,
//                   which needs exactly that velocity, runs here too:
//                   G_{k-1} = Ct(tracer_f_{k-1}, u_k), where tracer_f_0 is the
//                   input and tracer_f_{k-1} = B7 S7 G_{k-2} otherwise.
//                   k < n collides f_k into F_k; k = n writes f_n and the moments.
//   tracer launch   tracer_f_n = B7 S7 G_{n-1}, tracer = sum (0 at walls).
// Multiphase, also pull with post-collision buffers, two launches per step:
//   psi pass     gathers f_k = B S F_{k-1} (f_0 itself for k = 0), applies the
//                density floor of lbm_pipe.py:802-814 as a per-cell scale s
//                (kept for the next pass) and writes psi(rho) (psi_wall at walls)
//   collide pass gathers again, scales by s, adds the Shan-Chen force from
//                psi at x + e_q and collides with exact-difference forcing into F_k
// and a last launch writes the floored f_n and the outputs (tracer =
// (rho - rho_gas)/(rho_liquid - rho_gas) clipped to [0, 1]); tracer_f is
// passed through unchanged (lbm_pipe.py:826-832).
#include <hip/hip_runtime.h>

#include "../../include/lbm_c_api.h"
#include "lbm_common.hpp"
#include "lbm_host.hpp"

namespace mlbm {

struct Q7 {
  static constexpr int ea[7] = {0, 1, -1, 0, 0, 0, 0};
  static constexpr int eb[7] = {0, 0, 0, 1, -1, 0, 0};
  static constexpr int ec[7] = {0, 0, 0, 0, 0, 1, -1};
  static constexpr int opp[7] = {0, 2, 1, 4, 3, 6, 5};
  static constexpr int w8[7] = {2, 1, 1, 1, 1, 1, 1};  // 1/4, 1/8 (x6)
};

template <class T>
struct PipeArgs {
  const T* fsrc;
  T* fdst;
  const T* gsrc;
  T* gdst;
  const uint8_t* flags;
  T* psi;
  T* scale;
  T* rho;
  T* u;
  T* tracer;
  long long plane;
  int na, nb, nc, bpr;
  int nblocks, remap;  // remap: xcd_remap run length (0: identity)
  int gsrc_aligned;   // tracer_f_{k-1} is the input itself (k = 1)
  T tau, tau_tracer, strength, gravity;
  T G, rho_0, psi_wall, rho_liquid, rho_gas;
};

template <class T>
__device__ __forceinline__ T w19(int q) {
  return T(Lattice<19>::w36[q]) / T(36);
}

// feq_q = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u)   (lbm_pipe.py:144-167)
template <class T>
__device__ __forceinline__ T feq(int q, T rho, T ua, T ub, T uc, T usq) {
  using L = Lattice<19>;
  const T eu = edot<L, T>(q, ua, ub, uc);
  return w19<T>(q) * rho * (T(1) + T(3) * eu + T(4.5) * eu * eu - T(1.5) * usq);
}

template <class T>
__device__ __forceinline__ void cell_index(const PipeArgs<T>& a, int& ia, int& ib, int& ic,
                                           bool& valid) {
  const int lb = xcd_remap(blockIdx.x, a.nblocks, a.remap);
  int row;
  if (a.bpr == 0) {
    // 256-cell tiles of the flattened lattice
    const int cell = lb * 256 + (int)threadIdx.x;
    valid = cell < (int)a.plane;
    row = (int)((unsigned)cell / (unsigned)a.nc);
    ic = cell - row * a.nc;
  } else {
    row = lb / a.bpr;
    ic = (lb - row * a.bpr) * blockDim.x + threadIdx.x;
    valid = ic < a.nc;
  }
  ia = row / a.nb;
  ib = row - ia * a.nb;
}

// ---- single phase -----------------------------------------------------------
template <class T, int PHASE>
__global__ void __launch_bounds__(256) k_pipe(PipeArgs<T> a) {
  using L = Lattice<19>;
  int ia, ib, ic;
  bool valid;
  cell_index(a, ia, ib, ic, valid);
  if (!valid) return;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const long long P = a.plane;
  const uint8_t fl = a.flags[n.cell];
  const bool wall = fl & 1;
  T f[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    if (PHASE == 0) {
      f[q] = a.fsrc[q * P + n.cell];
    } else {
      // streamed slot q came from x - e_q; a wall cell takes slot opp(q),
      // which came from x + e_q (lbm_pipe.py:794-797)
      f[q] = wall ? a.fsrc[L::opp[q] * P + n.template at<L>(q, +1)]
                  : a.fsrc[q * P + n.template at<L>(q, -1)];
    }
  }
  T rho = 0, ma = 0, mb = 0, mc = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += f[q];
    if (L::ea[q]) ma += T(L::ea[q]) * f[q];
    if (L::eb[q]) mb += T(L::eb[q]) * f[q];
    if (L::ec[q]) mc += T(L::ec[q]) * f[q];
  }
  const T inv = T(1) / (rho > T(1e-10) ? rho : T(1e-10));
  const T ua = ma * inv, ub = mb * inv, uc = mc * inv;   // no force term (lbm_pipe.py:170-187)
  if (PHASE >= 1) {
    // outputs of the previous step, then that step's tracer collision
    const T oa = wall ? T(0) : ua, ob = wall ? T(0) : ub, oc = wall ? T(0) : uc;
    T g[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      if (a.gsrc_aligned) {
        g[q] = a.gsrc[q * P + n.cell];
      } else {
        const int dq[3] = {Q7::ea[q], Q7::eb[q], Q7::ec[q]};
        const int up = n.row[1 + dq[0]][1 + dq[1]] + n.col[1 + dq[2]];
        const int dn = n.row[1 - dq[0]][1 - dq[1]] + n.col[1 - dq[2]];
        g[q] = wall ? a.gsrc[Q7::opp[q] * P + up] : a.gsrc[q * P + dn];
      }
    }
    T c = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q) c += g[q];
    if (!wall) {
      // g_eq = w7 C (1 + e.u/cs2); no collision at walls (lbm_pipe.py:256-271, 838-843)
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        const T eu = edot<Q7, T>(q, oa, ob, oc);
        const T ge = T(Q7::w8[q]) / T(8) * c * (T(1) + T(3) * eu);
        g[q] = g[q] - (g[q] - ge) / a.tau_tracer;
      }
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) __builtin_nontemporal_store(g[q], a.gdst + q * P + n.cell);
    if (PHASE == 2) {
#pragma unroll
      for (int q = 0; q < 19; ++q) __builtin_nontemporal_store(f[q], a.fdst + q * P + n.cell);
      if (a.rho) a.rho[n.cell] = rho;
      if (a.u) {
        a.u[3 * (long long)n.cell + 0] = oa;
        a.u[3 * (long long)n.cell + 1] = ob;
        a.u[3 * (long long)n.cell + 2] = oc;
      }
      return;
    }
  }
  // BGK + Guo on the raw velocity (lbm_pipe.py:781-787, 218-253)
  const T Fa = (fl & 2) ? a.strength : T(0);
  const T Fc = wall ? T(0) : a.gravity;
  const T usq = ua * ua + ub * ub + uc * uc;
  const T uF = ua * Fa + uc * Fc;
  const T pref = T(1) - T(0.5) / a.tau;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const T eu = edot<L, T>(q, ua, ub, uc);
    const T eF = T(L::ea[q]) * Fa + T(L::ec[q]) * Fc;
    const T fe = feq<T>(q, rho, ua, ub, uc, usq);
    const T S = pref * w19<T>(q) * (T(3) * (eF - uF) + T(9) * eu * eF);
    __builtin_nontemporal_store(f[q] - (f[q] - fe) / a.tau + S, a.fdst + q * P + n.cell);
  }
}

// tracer_f_n = B7 S7 G_{n-1}; tracer = sum g, 0 at walls (lbm_pipe.py:844-853)
template <class T>
__global__ void __launch_bounds__(256) k_pipe_tracer(PipeArgs<T> a) {
  int ia, ib, ic;
  bool valid;
  cell_index(a, ia, ib, ic, valid);
  if (!valid) return;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const long long P = a.plane;
  const bool wall = a.flags[n.cell] & 1;
  T c = 0;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int dq[3] = {Q7::ea[q], Q7::eb[q], Q7::ec[q]};
    const int up = n.row[1 + dq[0]][1 + dq[1]] + n.col[1 + dq[2]];
    const int dn = n.row[1 - dq[0]][1 - dq[1]] + n.col[1 - dq[2]];
    const T g = wall ? a.gsrc[Q7::opp[q] * P + up] : a.gsrc[q * P + dn];
    __builtin_nontemporal_store(g, a.gdst + q * P + n.cell);
    c += g;
  }
  if (a.tracer) a.tracer[n.cell] = wall ? T(0) : c;
}

// ---- multiphase ---------------------------------------------------------------
template <class T>
__device__ __forceinline__ T psi_of(T rho, T rho_0) {
  return rho_0 * (T(1) - exp(-rho / rho_0));
}

// f_k(x) gathered from the post-collision buffer with the node swap at wall
// cells, as in the single-phase kernel
template <class T>
__device__ __forceinline__ void gather19(const T* src, long long P, const Nbr& n, bool wall,
                                         T (&f)[19]) {
  using L = Lattice<19>;
#pragma unroll
  for (int q = 0; q < 19; ++q)
    f[q] = wall ? src[L::opp[q] * P + n.template at<L>(q, +1)]
                : src[q * P + n.template at<L>(q, -1)];
}

// Density floor of lbm_pipe.py:802-814 as a per-cell scale: s = 0.01/rho
// where the streamed density is below 0.01, else 1. The state the next step
// sees is s * (B S F).
template <class T>
__device__ __forceinline__ T floor_scale(T rho) {
  return rho < T(0.01) ? (rho > T(0.01) ? rho : T(0.01)) / (rho > T(1e-20) ? rho : T(1e-20))
                       : T(1);
}

// psi pass of step k: psi(x) from the density of f_k (FIRST: f_0 read in
// place; otherwise gathered from F_{k-1} and floored, the scale kept for
// the collision pass).
template <class T, bool FIRST>
__global__ void __launch_bounds__(256) k_pipe_psi(PipeArgs<T> a) {
  int ia, ib, ic;
  bool valid;
  cell_index(a, ia, ib, ic, valid);
  if (!valid) return;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const long long P = a.plane;
  const bool wall = a.flags[n.cell] & 1;
  T f[19];
  if (FIRST) {
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = a.fsrc[q * P + n.cell];
  } else {
    gather19<T>(a.fsrc, P, n, wall, f);
  }
  T rho = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) rho += f[q];
  if (!FIRST) {
    const T s = floor_scale<T>(rho);
    a.scale[n.cell] = s;
    if (s != T(1)) {
      rho = 0;
#pragma unroll
      for (int q = 0; q < 19; ++q) rho += f[q] * s;
    }
  }
  a.psi[n.cell] = wall ? a.psi_wall : psi_of<T>(rho, a.rho_0);
}

// collision of step k with Shan-Chen + exact-difference forcing
// (lbm_pipe.py:770-787, 321-375), written in place (pull scheme)
template <class T, bool FIRST>
__global__ void __launch_bounds__(256) k_pipe_mp(PipeArgs<T> a) {
  using L = Lattice<19>;
  int ia, ib, ic;
  bool valid;
  cell_index(a, ia, ib, ic, valid);
  if (!valid) return;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const long long P = a.plane;
  const uint8_t fl = a.flags[n.cell];
  const bool wall = fl & 1;
  T f[19];
  if (FIRST) {
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = a.fsrc[q * P + n.cell];
  } else {
    gather19<T>(a.fsrc, P, n, wall, f);
    const T s = a.scale[n.cell];
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = s != T(1) ? f[q] * s : f[q];
  }
  T rho = 0, ma = 0, mb = 0, mc = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += f[q];
    if (L::ea[q]) ma += T(L::ea[q]) * f[q];
    if (L::eb[q]) mb += T(L::eb[q]) * f[q];
    if (L::ec[q]) mc += T(L::ec[q]) * f[q];
  }
  const T inv = T(1) / (rho > T(1e-10) ? rho : T(1e-10));
  const T ua = ma * inv, ub = mb * inv, uc = mc * inv;
  // F_sc = -G psi(x) sum_q w_q psi(x + e_q) e_q, zero at walls
  T ga = 0, gb = 0, gc = 0;
#pragma unroll
  for (int q = 1; q < 19; ++q) {
    const T ps = w19<T>(q) * a.psi[n.template at<L>(q, +1)];
    if (L::ea[q]) ga += T(L::ea[q]) * ps;
    if (L::eb[q]) gb += T(L::eb[q]) * ps;
    if (L::ec[q]) gc += T(L::ec[q]) * ps;
  }
  const T px = a.psi[n.cell];
  T Fa = wall ? T(0) : -a.G * px * ga;
  T Fb = wall ? T(0) : -a.G * px * gb;
  T Fc = wall ? T(0) : -a.G * px * gc;
  Fa += (fl & 2) ? a.strength : T(0);
  Fc += wall ? T(0) : a.gravity;
  // u_s = u + F / max(rho, 0.01), clamped to |u_s| <= 0.25
  const T rs = rho > T(0.01) ? rho : T(0.01);
  T sa = ua + Fa / rs, sb = ub + Fb / rs, sc = uc + Fc / rs;
  const T mag = sqrt(sa * sa + sb * sb + sc * sc);
  const T lim = T(0.25) / (mag > T(1e-10) ? mag : T(1e-10));
  const T scale = lim < T(1) ? lim : T(1);
  sa *= scale;
  sb *= scale;
  sc *= scale;
  const T usq = ua * ua + ub * ub + uc * uc;
  const T ssq = sa * sa + sb * sb + sc * sc;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const T fb = feq<T>(q, rho, ua, ub, uc, usq);
    const T fsh = feq<T>(q, rho, sa, sb, sc, ssq);
    __builtin_nontemporal_store(f[q] - (f[q] - fb) / a.tau + (fsh - fb), a.fdst + q * P + n.cell);
  }
}

// last launch: f_n = floor(B S F_{n-1}) and the outputs; velocity from the
// unscaled state, 0 at walls; tracer from the floored density
// (lbm_pipe.py:799-832)
template <class T>
__global__ void __launch_bounds__(256) k_pipe_mp_out(PipeArgs<T> a) {
  using L = Lattice<19>;
  int ia, ib, ic;
  bool valid;
  cell_index(a, ia, ib, ic, valid);
  if (!valid) return;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const long long P = a.plane;
  const bool wall = a.flags[n.cell] & 1;
  T f[19];
  gather19<T>(a.fsrc, P, n, wall, f);
  T rho = 0, ma = 0, mb = 0, mc = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    rho += f[q];
    if (L::ea[q]) ma += T(L::ea[q]) * f[q];
    if (L::eb[q]) mb += T(L::eb[q]) * f[q];
    if (L::ec[q]) mc += T(L::ec[q]) * f[q];
  }
  const T inv = T(1) / (rho > T(1e-10) ? rho : T(1e-10));
  const T s = floor_scale<T>(rho);
#pragma unroll
  for (int q = 0; q < 19; ++q)
    __builtin_nontemporal_store(s != T(1) ? f[q] * s : f[q], a.fdst + q * P + n.cell);
  const T rho_out = s != T(1) ? T(0.01) : rho;
  if (a.rho) a.rho[n.cell] = rho_out;
  if (a.u) {
    a.u[3 * (long long)n.cell + 0] = wall ? T(0) : ma * inv;
    a.u[3 * (long long)n.cell + 1] = wall ? T(0) : mb * inv;
    a.u[3 * (long long)n.cell + 2] = wall ? T(0) : mc * inv;
  }
  if (a.tracer) {
    T t = (rho_out - a.rho_gas) / (a.rho_liquid - a.rho_gas);
    t = t < T(0) ? T(0) : (t > T(1) ? T(1) : t);
    a.tracer[n.cell] = wall ? T(0) : t;
  }
}

// ---- host -------------------------------------------------------------------
struct PipeGeom {
  long long plane;
  int na, nb, nc, bx, bpr, nblocks;
};

inline int pipe_geom(const lbm_pipe_desc* d, PipeGeom* g) {
  if (!d) return fail(LBM_ERR_INVALID, "desc must be non-NULL");
  if (d->storage != LBM_F32 && d->storage != LBM_F64)
    return fail(LBM_ERR_UNSUPPORTED, "the pipe step stores f as f32 or f64");
  for (int i = 0; i < 3; ++i)
    if (d->dims[i] < 1) return fail(LBM_ERR_INVALID, "dims must be >= 1");
  if (!(d->tau > 0.5)) return fail(LBM_ERR_INVALID, "tau must be > 0.5");
  if (!d->multiphase && !(d->tau_tracer > 0.5))
    return fail(LBM_ERR_INVALID, "tau_tracer must be > 0.5");
  if (d->multiphase && !(d->rho_0 > 0.0)) return fail(LBM_ERR_INVALID, "rho_0 must be > 0");
  if (d->multiphase && !(d->rho_liquid > d->rho_gas))
    return fail(LBM_ERR_INVALID, "rho_liquid must be > rho_gas");
  const long long plane = d->dims[0] * d->dims[1] * d->dims[2];
  if (plane > 0x7fffffffLL || d->dims[0] * d->dims[1] > 0x7fffffffLL)
    return fail(LBM_ERR_INVALID, "grid too large (cells per population plane must fit int32)");
  g->plane = plane;
  g->na = (int)d->dims[0];
  g->nb = (int)d->dims[1];
  g->nc = (int)d->dims[2];
  // as the fluid kernels: row blocks of 256 threads when rows are a multiple
  // of 256 cells, else 256-cell tiles of the flattened grid (bpr 0)
  g->bx = 256;
  if (g->nc % 256 == 0) {
    g->bpr = g->nc / 256;
    g->nblocks = g->na * g->nb * g->bpr;
  } else {
    g->bpr = 0;
    g->nblocks = (int)((plane + 255) / 256);
  }
  return LBM_OK;
}

inline size_t pipe_work_bytes(const lbm_pipe_desc* d, const PipeGeom& g) {
  const size_t s = d->storage == LBM_F64 ? 8 : 4;
  // single phase: F ping-pong (2 x 19 planes) + G ping-pong (2 x 7 planes);
  // multiphase: F ping-pong (2 x 19 planes) + psi + floor scale
  const size_t planes = d->multiphase ? 40 : 52;
  return planes * (size_t)g.plane * s;
}

template <class T>
int pipe_run(hipStream_t s, const lbm_pipe_desc* d, const PipeGeom& g, const T* f_in, T* f_out,
             const T* g_in, T* g_out, const uint8_t* flags, T* work, T* rho, T* u, T* tracer,
             int nsteps) {
  PipeArgs<T> a = {};
  a.flags = flags;
  a.plane = g.plane;
  a.na = g.na;
  a.nb = g.nb;
  a.nc = g.nc;
  a.bpr = g.bpr;
  a.nblocks = g.nblocks;
  a.remap = (g.bpr > 0 && g.nblocks <= kAutoXcdMaxBlocks && pow2((long long)g.nb * g.nc))
                ? kAutoXcdGroup : 0;
  a.tau = T(d->tau);
  a.tau_tracer = T(d->tau_tracer);
  a.strength = T(d->propeller_strength);
  a.gravity = T(d->gravity);
  a.G = T(d->G);
  a.rho_0 = T(d->rho_0);
  a.psi_wall = T(d->rho_0 * (1.0 - exp(-d->rho_wall / d->rho_0)));
  a.rho_liquid = T(d->rho_liquid);
  a.rho_gas = T(d->rho_gas);
  const long long P = g.plane;
  const dim3 grid(g.nblocks), block(g.bx);
  if (!d->multiphase) {
    T* F[2] = {work, work + 19 * P};
    T* Gb[2] = {work + 38 * P, work + 45 * P};
    a.fsrc = f_in;
    a.fdst = F[0];
    hipLaunchKernelGGL((k_pipe<T, 0>), grid, block, 0, s, a);
    for (int k = 1; k <= nsteps; ++k) {
      a.fsrc = F[(k - 1) & 1];
      a.gsrc = k == 1 ? g_in : Gb[k & 1];      // G_{k-2}
      a.gsrc_aligned = k == 1;
      a.gdst = Gb[(k - 1) & 1];                // G_{k-1}
      if (k < nsteps) {
        a.fdst = F[k & 1];
        hipLaunchKernelGGL((k_pipe<T, 1>), grid, block, 0, s, a);
      } else {
        a.fdst = f_out;
        a.rho = rho;
        a.u = u;
        hipLaunchKernelGGL((k_pipe<T, 2>), grid, block, 0, s, a);
      }
    }
    a.gsrc = Gb[(nsteps - 1) & 1];
    a.gsrc_aligned = 0;
    a.gdst = g_out;
    a.tracer = tracer;
    hipLaunchKernelGGL((k_pipe_tracer<T>), grid, block, 0, s, a);
    return check_launch("pipe single-phase launch");
  }
  T* S[2] = {work, work + 19 * P};
  a.psi = work + 38 * P;
  a.scale = work + 39 * P;
  for (int k = 0; k < nsteps; ++k) {
    a.fsrc = k == 0 ? f_in : S[(k - 1) & 1];   // f_0, or F_{k-1}
    a.fdst = S[k & 1];                          // F_k
    if (k == 0) {
      hipLaunchKernelGGL((k_pipe_psi<T, true>), grid, block, 0, s, a);
      hipLaunchKernelGGL((k_pipe_mp<T, true>), grid, block, 0, s, a);
    } else {
      hipLaunchKernelGGL((k_pipe_psi<T, false>), grid, block, 0, s, a);
      hipLaunchKernelGGL((k_pipe_mp<T, false>), grid, block, 0, s, a);
    }
  }
  a.fsrc = S[(nsteps - 1) & 1];
  a.fdst = f_out;
  a.rho = rho;
  a.u = u;
  a.tracer = tracer;
  hipLaunchKernelGGL((k_pipe_mp_out<T>), grid, block, 0, s, a);
  if (g_out != g_in) {
    if (hipMemcpyAsync(g_out, g_in, 7 * (size_t)P * sizeof(T), hipMemcpyDeviceToDevice, s) !=
        hipSuccess)
      return fail(LBM_ERR_HIP, "tracer_f copy failed");
  }
  return check_launch("pipe multiphase launch");
}

}  // namespace mlbm

using namespace mlbm;

extern "C" {

size_t lbm_pipe_workspace_size(const lbm_pipe_desc* d) {
  PipeGeom g;
  if (pipe_geom(d, &g)) return 0;
  return pipe_work_bytes(d, g);
}

int lbm_pipe_step(void* stream, const lbm_pipe_desc* d, const void* f_in, void* f_out,
                  const void* g_in, void* g_out, const uint8_t* cell_flags, void* work,
                  void* rho, void* u, void* tracer, int nsteps) {
  g_last_error.clear();
  (void)hipGetLastError();
  PipeGeom g;
  int rc = pipe_geom(d, &g);
  if (rc) return rc;
  if (nsteps < 1) return fail(LBM_ERR_INVALID, "nsteps must be >= 1");
  if (!f_in || !f_out || !g_in || !g_out || !cell_flags || !work)
    return fail(LBM_ERR_INVALID, "f_in, f_out, g_in, g_out, cell_flags and work must be non-NULL");
  if (f_in == f_out) return fail(LBM_ERR_INVALID, "f_in and f_out must be distinct");
  if (!d->multiphase && g_in == g_out) return fail(LBM_ERR_INVALID, "g_in and g_out must be distinct");
  hipStream_t s = (hipStream_t)stream;
  if (d->storage == LBM_F64)
    return pipe_run<double>(s, d, g, (const double*)f_in, (double*)f_out, (const double*)g_in,
                            (double*)g_out, cell_flags, (double*)work, (double*)rho, (double*)u,
                            (double*)tracer, nsteps);
  return pipe_run<float>(s, d, g, (const float*)f_in, (float*)f_out, (const float*)g_in,
                         (float*)g_out, cell_flags, (float*)work, (float*)rho, (float*)u,
                         (float*)tracer, nsteps);
}

}  // extern "C"
