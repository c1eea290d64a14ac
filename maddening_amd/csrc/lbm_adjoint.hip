// Discrete adjoint of the LBM step (reverse mode through one
// collide -> stream -> bounce-back update of lbm.py:899-986).
//
// One step is f' = B S C(f): C the per-cell Guo-forced BGK collision, S the
// periodic shift, B the wall rule. Its transpose applied to a cotangent f'_bar:
//   g_bar = S^T B^T f'_bar   a gather from x + e_p (the forward pull reversed)
//   f_bar = C'(f)^T g_bar    local: (1 - 1/tau) g_bar_p + A + B . e_p, where
//                            A, B come from the cotangents of rho and u that
//                            feq and the Guo source pass back.
// The same pass accumulates dL/dtau(x) and dL/dF (per cell, or per workgroup
// partials for a uniform force). Gradient buffers are plain values (not
// deviations) in the compute type: float for f32 storage, double for f64.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lbm_c_api.h"
#include "lbm_cell.hpp"
#include "lbm_common.hpp"
#include "lbm_host.hpp"

namespace mlbm {

template <class C>
struct AdjArgs {
  const void* f;        // state f_n the step started from (storage type)
  const C* gin;         // dL/df_{n+1}
  C* gout;              // dL/df_n
  const uint8_t* mask;
  // halfway link words (lbm_links_build), read by the BB_HALF_LINKS kernels
  const uint32_t* links;
  const uint32_t* links_moving;
  const float* force_field;
  const float* tau_field;
  float* tau_bar;       // += dL/dtau(x)
  float* force_bar;     // += dL/dF(x), (*G, D)
  double* partial;      // += per-workgroup dL/dF for a uniform force, [nblocks][4]
  const float* rho_bar; // moments cotangents (moments adjoint only)
  const float* u_bar;
  const float* p_bar;
  long long plane;
  int na, nb, nc;
  int bpr, nblocks, remap;  // remap: xcd_remap run length G (0: identity)
  int out_swapped;      // write fbar_out with the slots of wall cells exchanged
  int in_swapped;       // fbar_in of the moments adjoint is in that form
  int out_faces;        // apply the transpose of the Zou-He faces to fbar_out
  // Zou-He faces, as in the forward arguments (lbm_cell.hpp FwdArgs)
  int zh_n;
  int zh_axis[2];
  int zh_side[2];
  int zh_slot[2];       // 0 inlet, 1 outlet: where dL/drho of face i goes
  C zh_rho[2];
  C inv_tau;
  C Fa, Fb, Fc;
  C lid_a, lid_b, lid_c;
};

// Per-workgroup partial sums: kPart doubles per workgroup, slots 0-2 dL/dF
// of a uniform force, slots 3-4 dL/drho of the inlet and outlet faces.
constexpr int kPart = 8;

// Sum NV values over the workgroup and add them to partial[kPart*lb + k]:
// an xor butterfly inside each 64-lane wave, then one barrier and a fixed-
// order sum of the NT/64 wave totals. The order is the same on every launch,
// so the partials are deterministic.
template <int NT, int NV>
__device__ __forceinline__ void block_accumulate(const double (&v)[NV], double* partial, int lb) {
  constexpr int NW = NT / 64;
  __shared__ double sh[NV][NW];
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = v[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += __shfl_xor(s[k], off, 64);
  }
  const int wave = (int)threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][wave] = s[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = sh[k][0];
#pragma unroll
      for (int w = 1; w < NW; ++w) t += sh[k][w];
      partial[kPart * (long long)lb + k] += t;
    }
  }
}

template <int NT>
__device__ __forceinline__ void accumulate_partials(double dFa, double dFb, double dFc,
                                                    double r0, double r1, bool faces,
                                                    double* partial, int lb) {
  if (faces) {
    const double v[5] = {dFa, dFb, dFc, r0, r1};
    block_accumulate<NT, 5>(v, partial, lb);
  } else {
    const double v[3] = {dFa, dFb, dFc};
    block_accumulate<NT, 3>(v, partial, lb);
  }
}

template <class L, class C>
__device__ __forceinline__ void add_force_bar(const AdjArgs<C>& a, int cell, C ga, C gb, C gc) {
  if (L::D == 3) {
    a.force_bar[3 * (long long)cell + 0] += float(ga);
    a.force_bar[3 * (long long)cell + 1] += float(gb);
    a.force_bar[3 * (long long)cell + 2] += float(gc);
  } else {
    a.force_bar[2 * (long long)cell + 0] += float(gb);
    a.force_bar[2 * (long long)cell + 1] += float(gc);
  }
}

template <class L, class C>
__device__ __forceinline__ void read_force(const AdjArgs<C>& a, int cell, C& Fa, C& Fb, C& Fc) {
  Fa = a.Fa;
  Fb = a.Fb;
  Fc = a.Fc;
  if (a.force_field) {
    if (L::D == 3) {
      Fa = C(a.force_field[3 * (long long)cell + 0]);
      Fb = C(a.force_field[3 * (long long)cell + 1]);
      Fc = C(a.force_field[3 * (long long)cell + 2]);
    } else {
      Fb = C(a.force_field[2 * (long long)cell + 0]);
      Fc = C(a.force_field[2 * (long long)cell + 1]);
    }
  }
}

// FPOST: `f` holds the post-collision state of the previous step and the
// step's starting state f_n is gathered from it exactly as the forward pull
// does; otherwise `f` holds f_n itself.
// GSW (reference wall rule only): gin holds the slots q, opp(q) of every wall
// cell exchanged, so the transposed node swap folds into a plain gather and
// no neighbour mask is read. a.out_swapped writes gout in the same form,
// with this cell's own mask.
// FACES: Zou-He faces are configured (a.zh_n > 0); the face-free kernels
// carry none of that code.
template <class L, class T, int BB, bool FPOST, bool GSW, bool FACES>
__global__ void __launch_bounds__(256) k_adjoint(AdjArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  const int lb = xcd_remap(blockIdx.x, a.nblocks, a.remap);
  int row, ic;
  bool valid;
  if (a.bpr == 0) {
    // 256-cell tiles of the flattened lattice (rows that row blocks would
    // leave misaligned or ragged, as in the forward)
    const int cell = lb * 256 + (int)threadIdx.x;
    valid = cell < (int)a.plane;
    row = (int)((unsigned)cell / (unsigned)a.nc);
    ic = cell - row * a.nc;
  } else {
    row = lb / a.bpr;
    ic = (lb - row * a.bpr) * blockDim.x + threadIdx.x;
    valid = ic < a.nc;
  }
  C dFa = 0, dFb = 0, dFc = 0;
  double rb0 = 0, rb1 = 0;
  if (valid) {
    const int ia = row / a.nb;
    const int ib = row - ia * a.nb;
    Nbr n;
    n.init(ia, ib, ic, a.na, a.nb, a.nc);
    const long long P = a.plane;
    uint32_t lk = 0;
    uint8_t self;
    if constexpr (BB == BB_HALF_LINKS) {
      lk = a.links[n.cell];
      self = (lk & kLinkSolid) ? uint8_t(1) : uint8_t(0);
    } else {
      self = a.mask ? a.mask[n.cell] : uint8_t(0);
    }
    // Under halfway bounce-back a solid cell is reset to rest every step:
    // nothing flows back into it, and its tau and force do not act. Its lanes
    // still run the whole update (on valid addresses) and only the results
    // are zeroed, so a wave mixing solid and fluid cells issues one load and
    // one store per slot instead of a branch per kind.
    const bool solid = (BB == BB_HALF || BB == BB_HALF_LINKS) && self;
    {
      // (cotangents are written with non-temporal stores, as the forward
      // writes its states: each is read once, by the next transposed step)
      // g_bar = S^T B^T f'_bar: the value the forward step moved out of
      // slot p at this cell went to x = cell + e_p (slot p, or slot opp(p) if
      // x is a wall under the reference rule); under halfway bounce-back a
      // population blocked by a wall came back to this cell in slot opp(p).
      // dL/dtau so far, read before any store (gout may alias it as far as
      // the compiler knows, so it could not move this load up by itself)
      const float tau_bar_in = a.tau_bar ? a.tau_bar[n.cell] : 0.0f;
      C gb[L::Q];
#pragma unroll
      for (int p = 0; p < L::Q; ++p) {
        const int x = n.template at<L>(p, +1);
        if (BB == BB_NONE || (BB == BB_REF && GSW)) {
          gb[p] = a.gin[p * P + x];
        } else if (BB == BB_REF) {
          gb[p] = a.gin[(a.mask[x] ? L::opp[p] : p) * P + x];
        } else if (BB == BB_HALF_LINKS) {
          // x = cell + e_p is a wall exactly when this cell's link opp(p) is
          // blocked (bit 0 is never set): one load, no neighbour mask
          const C* ptr = ((lk >> L::opp[p]) & 1u) ? a.gin + L::opp[p] * P + n.cell
                                                   : a.gin + p * P + x;
          gb[p] = *ptr;
        } else {
          // streamed slot loaded first (valid address either way), so only
          // blocked links wait for the neighbour's mask
          gb[p] = a.gin[p * P + x];
          if (p != 0 && a.mask[x] != 0) gb[p] = a.gin[L::opp[p] * P + n.cell];
        }
      }
      // forward quantities of this cell
      C d[L::Q];
      if constexpr (BB == BB_HALF_LINKS && FPOST)
        gather_links<L, T>(static_cast<const T*>(a.f), a.links_moving, P, n, lk, a.lid_a, a.lid_b,
                           a.lid_c, d);
      else
        gather_state<L, T, BB == BB_HALF_LINKS ? BB_HALF : BB, FPOST>(
            static_cast<const T*>(a.f), a.mask, P, n, self, a.lid_a, a.lid_b, a.lid_c, d);
      // the state the step started from includes the previous step's faces
      if (FPOST && FACES && !self) apply_faces<L, C>(a, ia, ib, ic, d);
      C Fa, Fb, Fc;
      read_force<L, C>(a, n.cell, Fa, Fb, Fc);
      C drho, rho, ua, ub, uc;
      moments<L, C>(d, Fa, Fb, Fc, drho, rho, ua, ub, uc);
      const C om = a.tau_field ? recip(C(a.tau_field[n.cell])) : a.inv_tau;
      const C c = C(1) - C(0.5) * om;
      const C usq = ua * ua + ub * ub + uc * uc;
      const C uF = ua * Fa + ub * Fb + uc * Fc;
      // Sums over q of gw_q = g_bar_q w_q times {1, e.u, (e.u)^2, e_a, (e.u) e_a,
      // (e.F) e_a}, plus tb = sum_q g_bar_q d f*_q / d omega with
      //   d f*_q / d omega = (feq_q - f_q) - 0.5 w [3 (e.F - u.F) + 9 (e.u)(e.F)].
      // Taken per pair (q, opp q): e.u, e.F and e flip sign between the two,
      // so only the sum s = gw_q + gw_m or the difference t = gw_q - gw_m of
      // the pair enters each term.
      C G0 = 0, Geu = 0, Geu2 = 0, Ga = 0, Gb = 0, Gc = 0;
      C Ha = 0, Hb = 0, Hc = 0;  // sum gw (e.F) e
      C Ka = 0, Kb = 0, Kc = 0;  // sum gw (e.u) e
      C tb = 0;
      const C Arest = fmadd(-rho, C(1.5) * usq, drho);  // even part of (feq - w)/w at e.u = 0
      const C uF3 = C(3) * uF;
      {
        const C w = weight<L, C>(0);
        const C gw = gb[0] * w;
        G0 = gw;
        // rest population: (feq - w)/w = Arest, source -3 u.F
        tb = fmadd(gw, fmadd(C(0.5), uF3, Arest), -gb[0] * d[0]);
      }
#pragma unroll
      for (int q = 1; q < L::Q; ++q) {
        if (L::opp[q] < q) continue;
        const int m = L::opp[q];
        const C w = weight<L, C>(q);
        const C gq = gb[q] * w, gm = gb[m] * w;
        const C sm = gq + gm, df = gq - gm;
        const C eu = edot<L, C>(q, ua, ub, uc);
        const C eF = edot<L, C>(q, Fa, Fb, Fc);
        G0 += sm;
        Geu = fmadd(df, eu, Geu);
        Geu2 = fmadd(sm * eu, eu, Geu2);
        const C seu = sm * eu, seF = sm * eF;
        if (L::ea[q]) { Ga += C(L::ea[q]) * df; Ka += C(L::ea[q]) * seu; Ha += C(L::ea[q]) * seF; }
        if (L::eb[q]) { Gb += C(L::eb[q]) * df; Kb += C(L::eb[q]) * seu; Hb += C(L::eb[q]) * seF; }
        if (L::ec[q]) { Gc += C(L::ec[q]) * df; Kc += C(L::ec[q]) * seu; Hc += C(L::ec[q]) * seF; }
        // even part: A - 0.5 E with A = Arest + 4.5 rho (e.u)^2, E = 9 (e.u)(e.F) - 3 u.F;
        // odd part:  B - 0.5 O with B = 3 rho e.u, O = 3 e.F
        const C even = fmadd(C(4.5) * rho * eu, eu, Arest) - C(0.5) * fmadd(C(9) * eu, eF, -uF3);
        const C odd = fmadd(C(3) * rho, eu, C(-1.5) * eF);
        tb = fmadd(sm, even, fmadd(df, odd, tb)) - fmadd(gb[q], d[q], gb[m] * d[m]);
      }
      // cotangents of rho (u held fixed) and u from feq (times omega) and the source (times c)
      const C rho_bar = om * (G0 * (C(1) - C(1.5) * usq) + C(3) * Geu + C(4.5) * Geu2);
      const C orho = om * rho;
      const C ubar_a = orho * (C(3) * Ga + C(9) * Ka - C(3) * ua * G0) + c * (C(9) * Ha - C(3) * Fa * G0);
      const C ubar_b = orho * (C(3) * Gb + C(9) * Kb - C(3) * ub * G0) + c * (C(9) * Hb - C(3) * Fb * G0);
      const C ubar_c = orho * (C(3) * Gc + C(9) * Kc - C(3) * uc * G0) + c * (C(9) * Hc - C(3) * Fc * G0);
      // u = (j + F/2) / rho: du/df_p = (e_p - u)/rho, drho/df_p = 1
      const C inv_rho = recip(rho > C(1e-10) ? rho : C(1e-10));
      const C Ba = ubar_a * inv_rho, Bb = ubar_b * inv_rho, Bc = ubar_c * inv_rho;
      const C A = rho_bar - (Ba * ua + Bb * ub + Bc * uc);
      const C keep = C(1) - om;
      // exchanged form: slot p carries the cotangent of opp(p), whose e is -e_p
      const bool xs = BB == BB_REF && a.out_swapped && self;
      if (FACES && a.out_faces && !self) {
        // f_n itself came out of the previous step's faces: pass the
        // cotangent back through their transpose
        C o[L::Q];
#pragma unroll
        for (int p = 0; p < L::Q; ++p)
          o[p] = keep * gb[p] + A + edot<L, C>(p, Ba, Bb, Bc);
        C rbf[2] = {C(0), C(0)};
        apply_faces_transpose<L, C>(a, ia, ib, ic, o, rbf);
        rb0 = solid ? 0.0 : double(rbf[0]);
        rb1 = solid ? 0.0 : double(rbf[1]);
#pragma unroll
        for (int p = 0; p < L::Q; ++p)
          __builtin_nontemporal_store(solid ? C(0) : o[p], a.gout + p * P + n.cell);
      } else {
#pragma unroll
        for (int p = 0; p < L::Q; ++p) {
          const C eB = edot<L, C>(p, Ba, Bb, Bc);
          const C g = xs ? gb[L::opp[p]] : gb[p];
          __builtin_nontemporal_store(solid ? C(0) : keep * g + A + (xs ? -eB : eB),
                                      a.gout + p * P + n.cell);
        }
      }
      if (a.tau_bar && !solid)
        __builtin_nontemporal_store(tau_bar_in + float(-om * om * tb), a.tau_bar + n.cell);
      // dL/dF: the source directly (u fixed) plus through u = (j + F/2)/rho
      if (!solid) {
        dFa = c * (C(3) * Ga - C(3) * ua * G0 + C(9) * Ka) + C(0.5) * Ba;
        dFb = c * (C(3) * Gb - C(3) * ub * G0 + C(9) * Kb) + C(0.5) * Bb;
        dFc = c * (C(3) * Gc - C(3) * uc * G0 + C(9) * Kc) + C(0.5) * Bc;
        if (a.force_bar) add_force_bar<L, C>(a, n.cell, dFa, dFb, dFc);
      }
    }
  }
  if (a.partial)
    accumulate_partials<256>(double(dFa), double(dFb), double(dFc), rb0, rb1,
                             FACES && a.out_faces, a.partial, lb);
}

// Cotangent of the output moments (lbm.py:970-976) pulled back onto f:
// f_bar_p += rho_bar + p_bar/3 + sum_a u_bar_a (e_pa - u_a)/rho (u_bar = 0
// at wall cells, where u is set to 0), and dL/dF += u_bar / (2 rho).
template <class L, class T>
__global__ void __launch_bounds__(256) k_moments_adjoint(AdjArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  const int lb = blockIdx.x;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  C dFa = 0, dFb = 0, dFc = 0;
  double rb0 = 0, rb1 = 0;
  if (cell < a.plane) {
    const long long P = a.plane;
    const T* f = static_cast<const T*>(a.f);
    C d[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; ++q) d[q] = St::template load<L>(f + q * P + cell, q);
    C Fa, Fb, Fc;
    read_force<L, C>(a, (int)cell, Fa, Fb, Fc);
    C drho, rho, ua, ub, uc;
    moments<L, C>(d, Fa, Fb, Fc, drho, rho, ua, ub, uc);
    C rb = a.rho_bar ? C(a.rho_bar[cell]) : C(0);
    if (a.p_bar) rb += C(a.p_bar[cell]) * C(1.0 / 3.0);
    C va = 0, vb = 0, vc = 0;
    const bool wall = a.mask && a.mask[cell];
    if (a.u_bar && !wall) {
      if (L::D == 3) {
        va = C(a.u_bar[3 * cell + 0]);
        vb = C(a.u_bar[3 * cell + 1]);
        vc = C(a.u_bar[3 * cell + 2]);
      } else {
        vb = C(a.u_bar[2 * cell + 0]);
        vc = C(a.u_bar[2 * cell + 1]);
      }
    }
    const C inv_rho = recip(rho > C(1e-10) ? rho : C(1e-10));
    const C Ba = va * inv_rho, Bb = vb * inv_rho, Bc = vc * inv_rho;
    const C A = rb - (Ba * ua + Bb * ub + Bc * uc);
    // at a wall cell u_bar is dropped (B = 0), so exchanging slots touches
    // only the incoming cotangent: read from opp(p) when exactly one of the
    // two buffers is in the exchanged form
    const bool cross = wall && (a.in_swapped != a.out_swapped);
    if (a.out_faces && !wall) {
      // f was written by the faces of the last step: transpose them too
      C o[L::Q];
#pragma unroll
      for (int p = 0; p < L::Q; ++p) {
        const C base = a.gin ? a.gin[p * P + cell] : C(0);
        o[p] = base + A + edot<L, C>(p, Ba, Bb, Bc);
      }
      const int ic = (int)(cell % a.nc);
      const int row = (int)(cell / a.nc);
      const int ib = row % a.nb, ia = row / a.nb;
      C rbf[2] = {C(0), C(0)};
      apply_faces_transpose<L, C>(a, ia, ib, ic, o, rbf);
      rb0 = double(rbf[0]);
      rb1 = double(rbf[1]);
#pragma unroll
      for (int p = 0; p < L::Q; ++p) __builtin_nontemporal_store(o[p], a.gout + p * P + cell);
    } else {
#pragma unroll
      for (int p = 0; p < L::Q; ++p) {
        const C base = a.gin ? a.gin[(cross ? L::opp[p] : p) * P + cell] : C(0);
        __builtin_nontemporal_store(base + A + edot<L, C>(p, Ba, Bb, Bc),
                                    a.gout + p * P + cell);
      }
    }
    dFa = C(0.5) * Ba;
    dFb = C(0.5) * Bb;
    dFc = C(0.5) * Bc;
    if (a.force_bar) add_force_bar<L, C>(a, (int)cell, dFa, dFb, dFc);
  }
  if (a.partial)
    accumulate_partials<256>(double(dFa), double(dFb), double(dFc), rb0, rb1,
                             a.zh_n > 0 && a.out_faces, a.partial, lb);
}

__global__ void __launch_bounds__(256) k_fill_zero(double* p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Fold the per-workgroup partials: MODE 0 gives dL/dF (slots 0-2; D2Q9's
// axes (b, c) are the reference's (0, 1)), MODE 1 dL/drho of the two faces
// (slots 3-4).
template <int D, int MODE>
__global__ void __launch_bounds__(256) k_partial_sum(const double* partial, long long nparts,
                                                     double* out) {
  __shared__ double sh[3][256];
  double s[3] = {0, 0, 0};
  const int off = MODE == 0 ? 0 : 3;
  const int nv = MODE == 0 ? 3 : 2;
  for (long long i = threadIdx.x; i < nparts; i += blockDim.x)
    for (int k = 0; k < nv; ++k) s[k] += partial[kPart * i + off + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (MODE == 1) {
      out[0] = sh[0][0];
      out[1] = sh[1][0];
    } else if (D == 3) {
      out[0] = sh[0][0];
      out[1] = sh[1][0];
      out[2] = sh[2][0];
    } else {
      out[0] = sh[1][0];
      out[1] = sh[2][0];
      out[2] = 0.0;
    }
  }
}

// ---- host ------------------------------------------------------------------
// Adjoint launches use 256-thread workgroups (the block reduction's width).
inline int adj_geom(const lbm_desc* d, Geom* g) {
  int rc = check_desc(d, g);
  if (rc) return rc;
  if (d->storage != LBM_F32 && d->storage != LBM_F64 && d->storage != LBM_F32_DEV)
    return fail(LBM_ERR_UNSUPPORTED, "the adjoint needs f32, f32dev or f64 storage");
  if (d->bb_mode == LBM_BB_HALFWAY && (d->flags & (LBM_FLAG_ADJ_IN_SWAPPED | LBM_FLAG_ADJ_OUT_SWAPPED)))
    return fail(LBM_ERR_UNSUPPORTED,
                "exchanged-slot cotangents are defined for the reference wall rule only");
  g->bx = 256;
  g->bpr = (g->nc + 255) / 256;
  g->nblocks = g->na * g->nb * g->bpr;
  // bpr 0 marks a launch over 256-cell tiles of the flattened lattice
  if (const int tiles = flat_tiles(d, *g)) {
    g->bpr = 0;
    g->nblocks = tiles;
  }
  return LBM_OK;
}

inline long long partial_slots(const Geom& g) {
  const long long cell_blocks = (g.plane + 255) / 256;
  return g.nblocks > cell_blocks ? g.nblocks : cell_blocks;
}

template <class C>
AdjArgs<C> adj_args(const lbm_desc* d, const Geom& g) {
  AdjArgs<C> a = {};
  a.plane = g.plane;
  a.na = g.na; a.nb = g.nb; a.nc = g.nc;
  a.bpr = g.bpr;
  a.nblocks = g.nblocks;
  // flattened tiles keep identity order unless a run length is asked for
  a.remap = (g.bpr == 0 && !(d->flags & LBM_FLAG_XCD_REMAP)) ? 0 : xcd_group(d, g);
  // the exchanged form only means something where walls swap slots
  a.in_swapped = (d->bb_mode == LBM_BB_REFERENCE && (d->flags & LBM_FLAG_ADJ_IN_SWAPPED)) ? 1 : 0;
  a.out_swapped = (d->bb_mode == LBM_BB_REFERENCE && (d->flags & LBM_FLAG_ADJ_OUT_SWAPPED)) ? 1 : 0;
  a.out_faces = (d->flags & LBM_FLAG_ADJ_OUT_FACES) ? 1 : 0;
  // faces in application order (inlet, outlet), axes as in the forward
  const int faces[2] = {d->inlet_face, d->outlet_face};
  const double rhos[2] = {d->inlet_rho, d->outlet_rho};
  a.zh_n = 0;
  for (int i = 0; i < 2; ++i) {
    if (faces[i] <= 0) continue;
    a.zh_axis[a.zh_n] = (faces[i] - 1) / 2 + (d->q == 19 ? 0 : 1);
    a.zh_side[a.zh_n] = (faces[i] - 1) % 2;
    a.zh_slot[a.zh_n] = i;
    a.zh_rho[a.zh_n] = C(rhos[i]);
    ++a.zh_n;
  }
  a.inv_tau = C(1.0 / d->tau);
  if (d->q == 19) {
    a.Fa = C(d->force[0]); a.Fb = C(d->force[1]); a.Fc = C(d->force[2]);
    a.lid_a = C(d->lid_velocity[0]); a.lid_b = C(d->lid_velocity[1]); a.lid_c = C(d->lid_velocity[2]);
  } else {
    a.Fa = C(0); a.Fb = C(d->force[0]); a.Fc = C(d->force[1]);
    a.lid_a = C(0); a.lid_b = C(d->lid_velocity[0]); a.lid_c = C(d->lid_velocity[1]);
  }
  return a;
}

template <class L, class T, bool FPOST, bool FACES>
int adjoint_launch(hipStream_t s, int bb, bool in_swapped, const Geom& g,
                   const AdjArgs<typename Storage<T>::C>& a) {
  switch (bb) {
    case LBM_BB_NONE:
      hipLaunchKernelGGL((k_adjoint<L, T, BB_NONE, FPOST, false, FACES>), dim3(g.nblocks),
                         dim3(256), 0, s, a);
      break;
    case LBM_BB_REFERENCE:
      if (in_swapped)
        hipLaunchKernelGGL((k_adjoint<L, T, BB_REF, FPOST, true, FACES>), dim3(g.nblocks),
                           dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((k_adjoint<L, T, BB_REF, FPOST, false, FACES>), dim3(g.nblocks),
                           dim3(256), 0, s, a);
      break;
    default:
      if (a.links)
        hipLaunchKernelGGL((k_adjoint<L, T, BB_HALF_LINKS, FPOST, false, FACES>), dim3(g.nblocks),
                           dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((k_adjoint<L, T, BB_HALF, FPOST, false, FACES>), dim3(g.nblocks),
                           dim3(256), 0, s, a);
  }
  return check_launch("k_adjoint launch");
}

template <class L, class T>
int adjoint_typed(hipStream_t s, const lbm_desc* d, const Geom& g, const void* f, int f_post,
                  const void* gin, void* gout, const uint8_t* mask, const uint32_t* links,
                  const float* ff, const float* tf, float* tau_bar, float* force_bar, void* work) {
  using C = typename Storage<T>::C;
  AdjArgs<C> a = adj_args<C>(d, g);
  a.f = f;
  a.gin = static_cast<const C*>(gin);
  a.gout = static_cast<C*>(gout);
  a.mask = d->bb_mode == LBM_BB_NONE ? nullptr : mask;
  if (d->bb_mode == LBM_BB_HALFWAY && links) {
    a.links = links;
    a.links_moving = links + g.plane;
  }
  a.force_field = ff;
  a.tau_field = tf;
  a.tau_bar = tau_bar;
  a.force_bar = force_bar;
  a.partial = static_cast<double*>(work);
  const bool sw = a.in_swapped != 0;
  if (a.zh_n > 0) {
    if (f_post) return adjoint_launch<L, T, true, true>(s, d->bb_mode, sw, g, a);
    return adjoint_launch<L, T, false, true>(s, d->bb_mode, sw, g, a);
  }
  if (f_post) return adjoint_launch<L, T, true, false>(s, d->bb_mode, sw, g, a);
  return adjoint_launch<L, T, false, false>(s, d->bb_mode, sw, g, a);
}

template <class L, class T>
int moments_adjoint_typed(hipStream_t s, const lbm_desc* d, const Geom& g, const void* f,
                          const uint8_t* mask, const float* ff, const float* rho_bar,
                          const float* u_bar, const float* p_bar, const void* gin, void* gout,
                          float* force_bar, void* work) {
  using C = typename Storage<T>::C;
  AdjArgs<C> a = adj_args<C>(d, g);
  a.f = f;
  a.mask = mask;
  a.force_field = ff;
  a.rho_bar = rho_bar;
  a.u_bar = u_bar;
  a.p_bar = p_bar;
  a.gin = static_cast<const C*>(gin);
  a.gout = static_cast<C*>(gout);
  a.force_bar = force_bar;
  a.partial = static_cast<double*>(work);
  const long long nb = (g.plane + 255) / 256;
  hipLaunchKernelGGL((k_moments_adjoint<L, T>), dim3((unsigned)nb), dim3(256), 0, s, a);
  return check_launch("k_moments_adjoint launch");
}

}  // namespace mlbm

using namespace mlbm;

extern "C" {

size_t lbm_adjoint_workspace_size(const lbm_desc* d) {
  Geom g;
  if (adj_geom(d, &g)) return 0;
  return (size_t)partial_slots(g) * kPart * sizeof(double);
}

int lbm_adjoint_begin(void* stream, const lbm_desc* d, void* work) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = adj_geom(d, &g);
  if (rc) return rc;
  if (!work) return fail(LBM_ERR_INVALID, "work must be non-NULL");
  const long long n = partial_slots(g) * kPart;
  hipLaunchKernelGGL(k_fill_zero, dim3(256), dim3(256), 0, (hipStream_t)stream,
                     static_cast<double*>(work), n);
  return check_launch("k_fill_zero launch");
}

int lbm_adjoint_step(void* stream, const lbm_desc* d, const void* f_n, int f_n_post,
                     const void* fbar_next, void* fbar_out, const uint8_t* mask,
                     const uint32_t* links, const float* force_field, const float* tau_field,
                     float* tau_bar, float* force_bar_field, void* work) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = adj_geom(d, &g);
  if (rc) return rc;
  if (!f_n || !fbar_next || !fbar_out)
    return fail(LBM_ERR_INVALID, "f_n, fbar_next and fbar_out must be non-NULL");
  if (fbar_next == fbar_out) return fail(LBM_ERR_INVALID, "fbar_next and fbar_out must be distinct");
  if (d->bb_mode != LBM_BB_NONE && !mask) return fail(LBM_ERR_INVALID, "bb_mode needs a wall mask");
  hipStream_t s = (hipStream_t)stream;
#define MLBM_ADJ(LAT, T)                                                                    \
  return adjoint_typed<LAT, T>(s, d, g, f_n, f_n_post, fbar_next, fbar_out, mask, links,    \
                               force_field, tau_field, tau_bar, force_bar_field, work)
  if (d->q == 19) {
    if (d->storage == LBM_F32_DEV) MLBM_ADJ(Lattice<19>, f32dev_t);
    if (d->storage == LBM_F32) MLBM_ADJ(Lattice<19>, float);
    MLBM_ADJ(Lattice<19>, double);
  }
  if (d->storage == LBM_F32_DEV) MLBM_ADJ(Lattice<9>, f32dev_t);
  if (d->storage == LBM_F32) MLBM_ADJ(Lattice<9>, float);
  MLBM_ADJ(Lattice<9>, double);
#undef MLBM_ADJ
}

int lbm_moments_adjoint(void* stream, const lbm_desc* d, const void* f, const uint8_t* mask,
                        const float* force_field, const float* rho_bar, const float* u_bar,
                        const float* p_bar, const void* fbar_in, void* fbar_out,
                        float* force_bar_field, void* work) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = adj_geom(d, &g);
  if (rc) return rc;
  if (!f || !fbar_out) return fail(LBM_ERR_INVALID, "f and fbar_out must be non-NULL");
  hipStream_t s = (hipStream_t)stream;
  if (d->q == 19) {
    if (d->storage == LBM_F32_DEV)
      return moments_adjoint_typed<Lattice<19>, f32dev_t>(s, d, g, f, mask, force_field, rho_bar, u_bar,
                                                         p_bar, fbar_in, fbar_out, force_bar_field, work);
    if (d->storage == LBM_F32)
      return moments_adjoint_typed<Lattice<19>, float>(s, d, g, f, mask, force_field, rho_bar, u_bar,
                                                      p_bar, fbar_in, fbar_out, force_bar_field, work);
    return moments_adjoint_typed<Lattice<19>, double>(s, d, g, f, mask, force_field, rho_bar, u_bar,
                                                     p_bar, fbar_in, fbar_out, force_bar_field, work);
  }
  if (d->storage == LBM_F32_DEV)
    return moments_adjoint_typed<Lattice<9>, f32dev_t>(s, d, g, f, mask, force_field, rho_bar, u_bar,
                                                      p_bar, fbar_in, fbar_out, force_bar_field, work);
  if (d->storage == LBM_F32)
    return moments_adjoint_typed<Lattice<9>, float>(s, d, g, f, mask, force_field, rho_bar, u_bar,
                                                   p_bar, fbar_in, fbar_out, force_bar_field, work);
  return moments_adjoint_typed<Lattice<9>, double>(s, d, g, f, mask, force_field, rho_bar, u_bar,
                                                  p_bar, fbar_in, fbar_out, force_bar_field, work);
}

int lbm_adjoint_end(void* stream, const lbm_desc* d, const void* work, double* force_bar_out) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = adj_geom(d, &g);
  if (rc) return rc;
  if (!work || !force_bar_out) return fail(LBM_ERR_INVALID, "work and force_bar_out must be non-NULL");
  if (d->q == 19)
    hipLaunchKernelGGL((k_partial_sum<3, 0>), dim3(1), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const double*>(work), partial_slots(g), force_bar_out);
  else
    hipLaunchKernelGGL((k_partial_sum<2, 0>), dim3(1), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const double*>(work), partial_slots(g), force_bar_out);
  return check_launch("k_partial_sum launch");
}

int lbm_adjoint_face_end(void* stream, const lbm_desc* d, const void* work, double* rho_bar_out) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = adj_geom(d, &g);
  if (rc) return rc;
  if (!work || !rho_bar_out) return fail(LBM_ERR_INVALID, "work and rho_bar_out must be non-NULL");
  hipLaunchKernelGGL((k_partial_sum<3, 1>), dim3(1), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const double*>(work), partial_slots(g), rho_bar_out);
  return check_launch("k_partial_sum launch");
}

}  // extern "C"
