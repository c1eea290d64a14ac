// Per-cell LBM update shared by the single-lattice kernels
// (lbm_forward.hip) and the batched kernel (lbm_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_common.hpp"
#include "lbm_host.hpp"

namespace mlbm {

template <class C>
struct FwdArgs {
  const void* src;
  void* dst;
  const uint8_t* mask;
  // per-cell link words for halfway bounce-back (k_links; read by the
  // BB_HALF_LINKS kernels): bit q set when the cell is fluid and x - e_q is a
  // wall, kLinkMoving when one of those walls moves, kLinkSolid for a wall
  // cell; links_moving[cell] has bit q set when x - e_q is a moving wall
  const uint32_t* links;
  const uint32_t* links_moving;
  const float* force_field;
  const float* tau_field;
  float* rho;
  float* u;
  float* p;
  long long plane;  // cells per population plane
  int na, nb, nc;
  int bpr;          // workgroups per grid row
  int nblocks;
  int remap;        // xcd_remap run length G (0: identity)
  // > 0: k_cell runs over flat_blocks tiles of 256 consecutive cells of the
  // flattened lattice instead of row blocks (flat_remap: their XCD runs)
  int flat_blocks;
  int flat_remap;
  // flattened tiles of 32- or 16-bit storage whose plane size is not a
  // multiple of 64 cells: stores go through LDS (k_cell_staged)
  int flat_staged;
  C inv_tau;
  C Fa, Fb, Fc;
  C lid_a, lid_b, lid_c;
  // Zou-He pressure faces applied after streaming, in order (inlet, outlet)
  int zh_n;
  int zh_axis[2];  // internal axis 0 (a), 1 (b), 2 (c)
  int zh_side[2];  // 0: first layer, 1: last layer
  C zh_rho[2];     // prescribed density P / cs2
};

constexpr uint32_t kLinkSolid = 1u << 31;
constexpr uint32_t kLinkMoving = 1u << 30;

// Zou-He pressure face at one fluid cell (lbm.py:425-510), on deviations.
// Directions pointing out of the domain through the face are unknown; the
// normal velocity follows the reference's form
//   u_n = +-(1 - (sum known + sum tangential) / rho_p)
// and each unknown population is rebuilt by non-equilibrium bounce-back,
// f_q = f_opp(q) + feq_q - feq_opp(q) = f_opp(q) + 6 w_q rho_p e_qn u_n.
// Only unknown slots are written, and they read known slots: in place is safe.
template <class L, class C>
__device__ __forceinline__ void zou_he(C (&d)[L::Q], int axis, int side, C rho_p) {
  C S = 0;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int e = axis == 0 ? L::ea[q] : (axis == 1 ? L::eb[q] : L::ec[q]);
    if (side == 0 ? e <= 0 : e >= 0) S += d[q] + weight<L, C>(q);
  }
  const C r = rho_p > C(1e-10) ? rho_p : C(1e-10);
  const C un = side == 0 ? C(1) - S / r : C(-1) + S / r;
  C nd[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int e = axis == 0 ? L::ea[q] : (axis == 1 ? L::eb[q] : L::ec[q]);
    const bool unknown = side == 0 ? e > 0 : e < 0;
    nd[q] = unknown ? d[L::opp[q]] + C(6) * weight<L, C>(q) * rho_p * C(e) * un : d[q];
  }
#pragma unroll
  for (int q = 0; q < L::Q; ++q) d[q] = nd[q];
}

// Apply the configured faces to a cell at (ia, ib, ic) that is not a wall.
// A is any argument block with the zh_* face fields and the extents.
template <class L, class C, class A>
__device__ __forceinline__ void apply_faces(const A& a, int ia, int ib, int ic, C (&d)[L::Q]) {
  for (int i = 0; i < a.zh_n; ++i) {
    const int ax = a.zh_axis[i];
    const int coord = ax == 0 ? ia : (ax == 1 ? ib : ic);
    const int n = ax == 0 ? a.na : (ax == 1 ? a.nb : a.nc);
    if (coord == (a.zh_side[i] == 0 ? 0 : n - 1)) zou_he<L, C>(d, ax, a.zh_side[i], a.zh_rho[i]);
  }
}

// Transpose of zou_he on a cotangent y (w.r.t. its output) at one cell; the
// face map is affine in the populations (its Jacobian depends only on rho_p):
//   unknown q:   new_q = d_opp(q) + 6 w_q rho_p e_qn u_n,
//                u_n = s (1 - S / r), s = +1 (first layer) or -1, r = max(rho_p, 1e-10),
//                S = sum of f over known and tangential slots,
// so d new_q / d d_p = [p == opp q] - 6 s w_q e_qn rho_p / r for every known or
// tangential p, unknown slots are overwritten (no dependence on their old
// value), and d new_q / d rho_p = 6 s w_q e_qn. Returns y mapped to the
// cotangent w.r.t. the face map's input, adds dL/drho_p to rho_bar.
template <class L, class C>
__device__ __forceinline__ void zou_he_transpose(C (&y)[L::Q], int axis, int side, C rho_p,
                                                 C& rho_bar) {
  const C r = rho_p > C(1e-10) ? rho_p : C(1e-10);
  const C sg = side == 0 ? C(1) : C(-1);
  C c = 0, rb = 0;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int e = axis == 0 ? L::ea[q] : (axis == 1 ? L::eb[q] : L::ec[q]);
    const bool unknown = side == 0 ? e > 0 : e < 0;
    if (unknown) {
      const C k6 = C(6) * weight<L, C>(q) * C(e);
      c = fmadd(y[q], -sg * k6 * rho_p / r, c);
      rb = fmadd(y[q], sg * k6, rb);
    }
  }
  C x[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int e = axis == 0 ? L::ea[q] : (axis == 1 ? L::eb[q] : L::ec[q]);
    const bool unknown = side == 0 ? e > 0 : e < 0;
    x[q] = unknown ? C(0) : y[q] + c;
  }
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int e = axis == 0 ? L::ea[q] : (axis == 1 ? L::eb[q] : L::ec[q]);
    const bool unknown = side == 0 ? e > 0 : e < 0;
    if (unknown) x[L::opp[q]] += y[q];
  }
#pragma unroll
  for (int q = 0; q < L::Q; ++q) y[q] = x[q];
  rho_bar += rb;
}

// Transpose of apply_faces (faces in reverse order); rho_bar[a.zh_slot[i]]
// += dL/drho_p of face i (slot 0 inlet, 1 outlet).
template <class L, class C, class A>
__device__ __forceinline__ void apply_faces_transpose(const A& a, int ia, int ib, int ic,
                                                      C (&y)[L::Q], C (&rho_bar)[2]) {
  for (int i = a.zh_n - 1; i >= 0; --i) {
    const int ax = a.zh_axis[i];
    const int coord = ax == 0 ? ia : (ax == 1 ? ib : ic);
    const int n = ax == 0 ? a.na : (ax == 1 ? a.nb : a.nc);
    if (coord == (a.zh_side[i] == 0 ? 0 : n - 1))
      zou_he_transpose<L, C>(y, ax, a.zh_side[i], a.zh_rho[i], rho_bar[a.zh_slot[i]]);
  }
}

template <class L, class C>
__device__ __forceinline__ void load_force(const FwdArgs<C>& a, int cell, C& Fa,
                                           C& Fb, C& Fc) {
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

template <class L, class C>
__device__ __forceinline__ void write_moments(const FwdArgs<C>& a, int cell,
                                              bool wall, C rho, C ua, C ub,
                                              C uc) {
  if (a.rho) a.rho[cell] = float(rho);
  if (a.p) a.p[cell] = float(rho * C(1.0 / 3.0));
  if (a.u) {
    if (wall) ua = ub = uc = C(0);
    if (L::D == 3) {
      a.u[3 * (long long)cell + 0] = float(ua);
      a.u[3 * (long long)cell + 1] = float(ub);
      a.u[3 * (long long)cell + 2] = float(uc);
    } else {
      a.u[2 * (long long)cell + 0] = float(ub);
      a.u[2 * (long long)cell + 1] = float(uc);
    }
  }
}

// 2 w_q (e_q . u_wall) / cs2: momentum a moving wall adds to the population
// it reflects into direction q (halfway bounce-back, wall density 1).
template <class L, class C>
__device__ __forceinline__ C lid_term(const FwdArgs<C>& a, int q) {
  return C(6) * weight<L, C>(q) *
         edot<L, C>(q, a.lid_a, a.lid_b, a.lid_c);
}

// The populations a cell holds after streaming and bounce-back, read from a
// post-collision state `src` (PULL), or the populations stored at the cell
// (!PULL). Returned as deviations d_q = f_q - w_q.
template <class L, class T, int BB, bool PULL>
__device__ __forceinline__ void gather_state(const T* src, const uint8_t* mask, long long P,
                                             const Nbr& n, uint8_t self,
                                             typename Storage<T>::C lid_a,
                                             typename Storage<T>::C lid_b,
                                             typename Storage<T>::C lid_c,
                                             typename Storage<T>::C (&d)[L::Q]) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  if (BB == BB_HALF && PULL && self) {
    // solid cell under halfway bounce-back: no fluid reads it; park it at rest
#pragma unroll
    for (int q = 0; q < L::Q; ++q) d[q] = C(0);
    return;
  }
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    if (!PULL) {
      d[q] = St::template load<L>(src + q * P + n.cell, q);
    } else if (BB == BB_REF) {
      // reference bounce-back (lbm.py:944-948): at a wall cell slot q takes
      // the value streamed into slot opp(q), which came from x + e_q.
      const T* ptr = self ? src + L::opp[q] * P + n.template at<L>(q, +1)
                          : src + q * P + n.template at<L>(q, -1);
      d[q] = St::template load<L>(ptr, q);
    } else if (BB == BB_HALF) {
      // the streamed value is loaded whatever the neighbour is (its address
      // is valid either way), so the common fluid-fluid link does not wait
      // for the mask; a blocked link then takes the cell's own reflected value
      const int s = n.template at<L>(q, -1);
      d[q] = St::template load<L>(src + q * P + s, q);
      const uint8_t m = (q == 0) ? uint8_t(0) : mask[s];
      if (m) {
        d[q] = St::template load<L>(src + L::opp[q] * P + n.cell, q);
        // 2 w_q (e_q . u_wall) / cs2: momentum a moving wall (mask 2) adds to
        // the population it reflects into direction q (wall density 1)
        if (m == 2)
          d[q] += C(6) * weight<L, C>(q) *
                  edot<L, C>(q, lid_a, lid_b, lid_c);
      }
    } else {
      d[q] = St::template load<L>(src + q * P + n.template at<L>(q, -1), q);
    }
  }
}

// Halfway bounce-back gather from a link word (see FwdArgs::links): the same
// values as gather_state<L, T, BB_HALF, true>, with one aligned 4-byte read
// in place of the cell's own and its 18 neighbours' mask bytes, and one load
// per direction; a cell next to a moving wall also reads its moving-link
// word, and no neighbour index is needed.
template <class L, class T>
__device__ __forceinline__ void gather_links(const T* src, const uint32_t* links_moving,
                                             long long P, const Nbr& n, uint32_t lk,
                                             typename Storage<T>::C lid_a,
                                             typename Storage<T>::C lid_b,
                                             typename Storage<T>::C lid_c,
                                             typename Storage<T>::C (&d)[L::Q]) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  if (lk & kLinkSolid) {
#pragma unroll
    for (int q = 0; q < L::Q; ++q) d[q] = C(0);
    return;
  }
  // one load per direction: the streamed value, or on a blocked link the
  // cell's own reflected value (no load is spent on the wall's population)
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const bool blocked = (lk >> q) & 1u;  // bit 0 is never set
    const T* ptr = blocked ? src + L::opp[q] * P + n.cell : src + q * P + n.template at<L>(q, -1);
    d[q] = St::template load<L>(ptr, q);
  }
  if (lk & kLinkMoving) {
    // 2 w_q (e_q . u_wall) / cs2 on links to a moving wall (density 1)
    const uint32_t mv = links_moving[n.cell];
#pragma unroll
    for (int q = 1; q < L::Q; ++q)
      if ((mv >> q) & 1u)
        d[q] += C(6) * weight<L, C>(q) *
                edot<L, C>(q, lid_a, lid_b, lid_c);
  }
}

// The cell's update up to the store: d receives the state to write (after
// the collision when COLLIDE); returns the cell's index.
// PULL: gather (stream + bounce-back) instead of an aligned read.
// COLLIDE: apply the collision after the read.
template <class L, class T, int BB, bool PULL, bool COLLIDE>
__device__ __forceinline__ int cell_compute(const FwdArgs<typename Storage<T>::C>& a, int row,
                                            int ic, typename Storage<T>::C (&d)[L::Q]) {
  using C = typename Storage<T>::C;
  const int ia = row / a.nb;
  const int ib = row - ia * a.nb;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const T* src = static_cast<const T*>(a.src);
  const long long P = a.plane;
  uint8_t self;
  if constexpr (BB == BB_HALF_LINKS && PULL) {
    const uint32_t lk = a.links[n.cell];
    self = (lk & kLinkSolid) ? uint8_t(1) : uint8_t(0);
    gather_links<L, T>(src, a.links_moving, P, n, lk, a.lid_a, a.lid_b, a.lid_c, d);
  } else {
    self = a.mask ? a.mask[n.cell] : uint8_t(0);
    gather_state<L, T, BB == BB_HALF_LINKS ? BB_HALF : BB, PULL>(src, a.mask, P, n, self, a.lid_a,
                                                                 a.lid_b, a.lid_c, d);
  }
  if (PULL && a.zh_n && !self) apply_faces<L, C>(a, ia, ib, ic, d);

  C Fa, Fb, Fc;
  load_force<L, C>(a, n.cell, Fa, Fb, Fc);
  C drho, rho, ua, ub, uc;
  moments<L, C>(d, Fa, Fb, Fc, drho, rho, ua, ub, uc);
  if (a.rho || a.u || a.p) write_moments<L, C>(a, n.cell, self != 0, rho, ua, ub, uc);
  if (COLLIDE) {
    const C inv_tau = a.tau_field ? recip(C(a.tau_field[n.cell])) : a.inv_tau;
    collide_for<L, T, C>(d, inv_tau, Fa, Fb, Fc, drho, rho, ua, ub, uc);
  }
  return n.cell;
}

// cell_compute, then the store (STORE false: the moments-only pass).
template <class L, class T, int BB, bool PULL, bool COLLIDE, bool STORE>
__device__ __forceinline__ void cell_update(const FwdArgs<typename Storage<T>::C>& a, int row,
                                            int ic) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  C d[L::Q];
  const int cell = cell_compute<L, T, BB, PULL, COLLIDE>(a, row, ic, d);
  if (STORE) {
    T* dst = static_cast<T*>(a.dst);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) St::template store<L>(dst + q * a.plane + cell, q, d[q]);
  }
}

template <class L, class T, int BB, bool PULL, bool COLLIDE, bool STORE>
__global__ void __launch_bounds__(256)
    k_cell(FwdArgs<typename Storage<T>::C> a) {
  int row, ic;
  if (a.flat_blocks > 0) {
    // 256-thread tiles of the flattened lattice: every store of a wave is
    // one aligned 256-byte run whenever the plane size is a multiple of 64
    const int lb = xcd_remap(blockIdx.x, a.flat_blocks, a.flat_remap);
    const int cell = lb * 256 + (int)threadIdx.x;
    if (cell >= (int)a.plane) return;
    row = (int)((unsigned)cell / (unsigned)a.nc);
    ic = cell - row * a.nc;
  } else {
    const int lb = xcd_remap(blockIdx.x, a.nblocks, a.remap);
    row = lb / a.bpr;
    ic = (lb - row * a.bpr) * blockDim.x + threadIdx.x;
    if (ic >= a.nc) return;
  }
  // one inlined update for both launch shapes keeps the register count of
  // a single copy
  cell_update<L, T, BB, PULL, COLLIDE, STORE>(a, row, ic);
}

// Flattened 256-cell tiles of a lattice whose plane size is not a multiple
// of 64 cells. Storing straight from registers, every wave's 64-element run
// would start part-way into a 64-element segment of each population plane
// (plane q starts at q * plane). The tile's results go through LDS instead,
// and each wave stores one 64-element run aligned in the destination plane;
// only the two runs at the tile's ends stay partial (they share segments
// with the neighbouring tiles). 32- and 16-bit storage only (19.5 KB of LDS
// for D3Q19 float).
template <class L, class T, int BB, bool PULL, bool COLLIDE>
__global__ void __launch_bounds__(256)
    k_cell_staged(FwdArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  __shared__ T stage[L::Q][256];
  const int lb = xcd_remap(blockIdx.x, a.flat_blocks, a.flat_remap);
  const int t0 = lb * 256;
  const int plane = (int)a.plane;
  const int n = plane - t0 < 256 ? plane - t0 : 256;  // cells in this tile
  const int tid = (int)threadIdx.x;
  if (tid < n) {
    const int cell = t0 + tid;
    const int row = (int)((unsigned)cell / (unsigned)a.nc);
    C d[L::Q];
    cell_compute<L, T, BB, PULL, COLLIDE>(a, row, cell - row * a.nc, d);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) stage[q][tid] = St::template enc<L>(q, d[q]);
  }
  __syncthreads();
  T* dst = static_cast<T*>(a.dst);
  const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    // element t of the tile goes to dst[base + t]; the first element at a
    // multiple of 64 is t = lead. Wave w stores t in [lead + 64 (w - 1),
    // lead + 64 w), wave 0 also [lead + 192, 256): every t in [0, n) once.
    const long long base = (long long)q * plane + t0;
    const int lead = (int)((64 - (base & 63)) & 63);
    int t = lead - 64 + 64 * wave + lane;
    if (t >= 0 && t < n) St::put(dst + base + t, stage[q][t]);
    if (wave == 0) {
      t = lead + 192 + lane;
      if (t < n) St::put(dst + base + t, stage[q][t]);
    }
  }
}

// Link words for halfway bounce-back (FwdArgs::links, links_moving), built
// at the start of each lbm_step call or batch run (the mask is an input of
// each): reads the cell's mask byte and its neighbours' (from cache), writes
// 8 bytes per cell.
template <class L>
__global__ void __launch_bounds__(256)
    k_links(const uint8_t* __restrict__ mask, uint32_t* __restrict__ links,
            uint32_t* __restrict__ links_moving, int na, int nb, int nc, int bpr) {
  const int row = blockIdx.x / bpr;
  const int ic = (blockIdx.x - row * bpr) * blockDim.x + threadIdx.x;
  if (ic >= nc) return;
  const int ia = row / nb;
  const int ib = row - ia * nb;
  Nbr n;
  n.init(ia, ib, ic, na, nb, nc);
  uint32_t w = 0, mv = 0;
  if (mask[n.cell]) {
    w = kLinkSolid;
  } else {
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
      const uint8_t m = mask[n.template at<L>(q, -1)];
      if (m) w |= 1u << q;
      if (m == 2) mv |= 1u << q;
    }
    if (mv) w |= kLinkMoving;
  }
  links[n.cell] = w;
  links_moving[n.cell] = mv;
}

}  // namespace mlbm
