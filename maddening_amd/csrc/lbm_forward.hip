// Forward kernels of the LBM step and the C ABI entry points that drive them.
//
// One thread owns one cell; a workgroup owns `block_x` consecutive cells of
// one grid row along the contiguous axis. Every kernel reads each population
// of its cell once and writes it once, as 19 (or 9) separate coalesced
// streams, one per population plane.
//
// Two ways to chain steps, both reading and writing each population exactly
// once per step:
//  * pull: the state kept between kernels is post-collision. One fused kernel
//    gathers f(x) = g(x - e) (the stream, with bounce-back), which is the
//    reference's state f after a step, collides it and writes g' at x.
//    A step sequence is collide-only, (n-1) fused kernels, then a final
//    stream kernel that also emits the moments.
//  * push: the state kept between kernels is the reference's f. One kernel
//    collides f(x) and scatters f*(x) to x + e (or to the bounce-back slot).
//    Moments come from a separate read-only pass at the end.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/lbm_c_api.h"
#include "lbm_common.hpp"
#include "lbm_host.hpp"
#include "lbm_cell.hpp"
#include "lbm_pair.hpp"
#include "lbm_tri.hpp"

namespace mlbm {


// Push: aligned read of the reference state f(x), collision, scatter of f*(x)
// to x + e_q. Under reference bounce-back a value entering a wall cell lands
// in the opposite slot (lbm.py:944-948); under halfway bounce-back a value
// heading into a wall returns to x in the opposite slot.
template <class L, class T, int BB>
__global__ void __launch_bounds__(256) k_push(FwdArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  const int lb = xcd_remap(blockIdx.x, a.nblocks, a.remap);
  const int row = lb / a.bpr;
  const int ic = (lb - row * a.bpr) * blockDim.x + threadIdx.x;
  if (ic >= a.nc) return;
  const int ia = row / a.nb;
  const int ib = row - ia * a.nb;
  Nbr n;
  n.init(ia, ib, ic, a.na, a.nb, a.nc);
  const T* src = static_cast<const T*>(a.src);
  T* dst = static_cast<T*>(a.dst);
  const long long P = a.plane;

  if (BB == BB_HALF && a.mask[n.cell]) {
#pragma unroll
    for (int q = 0; q < L::Q; ++q) St::template store<L>(dst + q * P + n.cell, q, C(0));
    return;
  }
  C d[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) d[q] = St::template load<L>(src + q * P + n.cell, q);
  C Fa, Fb, Fc;
  load_force<L, C>(a, n.cell, Fa, Fb, Fc);
  C drho, rho, ua, ub, uc;
  moments<L, C>(d, Fa, Fb, Fc, drho, rho, ua, ub, uc);
  const C inv_tau = a.tau_field ? recip(C(a.tau_field[n.cell])) : a.inv_tau;
  collide_for<L, T, C>(d, inv_tau, Fa, Fb, Fc, drho, rho, ua, ub, uc);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const int t = n.template at<L>(q, +1);
    if (BB == BB_NONE) {
      St::template store<L>(dst + q * P + t, q, d[q]);
    } else if (BB == BB_REF) {
      T* ptr = a.mask[t] ? dst + L::opp[q] * P + t : dst + q * P + t;
      St::template store<L>(ptr, q, d[q]);
    } else {
      const uint8_t m = (q == 0) ? uint8_t(0) : a.mask[t];
      if (m) {
        St::template store<L>(dst + L::opp[q] * P + n.cell, q,
                              d[q] - (m == 2 ? lid_term<L, C>(a, q) : C(0)));
      } else {
        St::template store<L>(dst + q * P + t, q, d[q]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Row-per-wave variant for periodic lattices whose contiguous axis is a
// multiple of 64*V cells. A wave owns 64*V consecutive cells of one row (lane
// l holds cells V*l .. V*l+V-1), so every population plane is read and written
// with V-wide vector accesses, and the row offsets of the 9 (a, b)
// neighbours are wave-uniform. The +-1 shift along the row is a lane rotation
// (DPP wave_ror / wave_rol); when a row spans several waves, the one value
// that crosses a wave boundary comes from a single-lane load.
// ---------------------------------------------------------------------------
template <class T, int V> struct VecT { typedef T type __attribute__((ext_vector_type(V))); };
template <class T> struct VecT<T, 1> { typedef T type; };

template <class T, int V>
__device__ __forceinline__ void vload(const T* p, T (&out)[V]) {
  using VT = typename VecT<T, V>::type;
  const VT v = *reinterpret_cast<const VT*>(p);
  if constexpr (V == 1) {
    out[0] = v;
  } else {
#pragma unroll
    for (int k = 0; k < V; ++k) out[k] = v[k];
  }
}

template <class T, int V>
__device__ __forceinline__ void vstore_nt(T* p, const T (&in)[V]) {
  using VT = typename VecT<T, V>::type;
  VT v;
  if constexpr (V == 1) {
    v = in[0];
  } else {
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = in[k];
  }
  __builtin_nontemporal_store(v, reinterpret_cast<VT*>(p));
}

// lane i receives lane i-1 (lane 0 receives lane 63): DPP wave_ror:1
__device__ __forceinline__ float lane_from_prev(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x13C, 0xF, 0xF, false));
}
// lane i receives lane i+1 (lane 63 receives lane 0): DPP wave_rol:1
__device__ __forceinline__ float lane_from_next(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_from_prev(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x13C, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x13C, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_from_next(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x134, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// HASF = false: the launch has no force (uniform zero, no field), so the
// collision leaves out its exact-zero force terms (same values).
template <class L, class T, int V, bool PULL, bool COLLIDE, bool HASF = true>
__global__ void __launch_bounds__(256) k_row(FwdArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  // float16 storage: deviations in stored units (exact, fewer multiplies)
  constexpr bool SCALED = ScaleOf<T>::v != 0;
  constexpr int LS = ScaleOf<T>::v;
  constexpr int SEG = 64 * V;
  const int lane = threadIdx.x & 63;
  const int segs = a.nc / SEG;  // wave segments per row
  // the wave's index is the same in every lane: taken from the first lane,
  // the row arithmetic below stays scalar and each load or store is one
  // scalar row base plus the lane's offset
  const long long wave = (long long)blockIdx.x * (blockDim.x / 64) +
                         __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long long nrows = (long long)a.na * a.nb;
  if (wave >= nrows * segs) return;  // whole wave leaves together
  const int row = (int)(wave / segs);
  const int seg = (int)(wave - (long long)row * segs);
  const int ia = row / a.nb;
  const int ib = row - ia * a.nb;
  const int A[3] = {ia == 0 ? a.na - 1 : ia - 1, ia, ia == a.na - 1 ? 0 : ia + 1};
  const int B[3] = {ib == 0 ? a.nb - 1 : ib - 1, ib, ib == a.nb - 1 ? 0 : ib + 1};
  const int c0 = seg * SEG;       // first cell of this wave's segment
  const int cl = c0 + lane * V;   // first cell of this lane
  // neighbour segment edges along the row (periodic)
  const int c_before = c0 == 0 ? a.nc - 1 : c0 - 1;
  const int c_after = c0 + SEG == a.nc ? 0 : c0 + SEG;
  const long long P = a.plane;
  const T* src = static_cast<const T*>(a.src);
  const uint32_t lofs = (uint32_t)lane * (uint32_t)(V * sizeof(T));

  C d[L::Q][V];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    // pull from x - e_q: source row (ia - ea, ib - eb)
    const int sr = PULL ? (A[1 - L::ea[q]] * a.nb + B[1 - L::eb[q]]) : row;
    const T* rowp = src + q * P + (long long)sr * a.nc;
    T raw[V];
    // scalar segment base plus the lane's 32-bit byte offset
    vload<T, V>(reinterpret_cast<const T*>(reinterpret_cast<const char*>(rowp + c0) + lofs), raw);
    C v[V];
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = load_u<L, T, SCALED>(&raw[k], q);
    if (!PULL || L::ec[q] == 0) {
#pragma unroll
      for (int k = 0; k < V; ++k) d[q][k] = v[k];
    } else if (L::ec[q] > 0) {
      // cell c takes c - 1: lane's first cell takes the previous lane's last;
      // lane 0 takes the cell before the segment (a wave-uniform load, no branch)
      const C before = load_u<L, T, SCALED>(rowp + c_before, q);
      const C edge = lane_from_prev(v[V - 1]);
      d[q][0] = lane == 0 ? before : edge;
#pragma unroll
      for (int k = 1; k < V; ++k) d[q][k] = v[k - 1];
    } else {
      // cell c takes c + 1: lane's last cell takes the next lane's first
      const C after = load_u<L, T, SCALED>(rowp + c_after, q);
      const C edge = lane_from_next(v[0]);
#pragma unroll
      for (int k = 0; k < V - 1; ++k) d[q][k] = v[k + 1];
      d[q][V - 1] = lane == 63 ? after : edge;
    }
  }

  const long long seg0 = (long long)row * a.nc + c0;  // first cell of the segment
  const long long cell0 = seg0 + lane * V;
  T out[L::Q][V];
#pragma unroll
  for (int k = 0; k < V; ++k) {
    C dk[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; ++q) dk[q] = d[q][k];
    const int cell = (int)(cell0 + k);
    C Fa = C(0), Fb = C(0), Fc = C(0);
    if constexpr (HASF) load_force<L, C>(a, cell, Fa, Fb, Fc);
    C drho, rho, ua, ub, uc;
    moments<L, C, HASF, LS>(dk, Fa, Fb, Fc, drho, rho, ua, ub, uc);
    if (a.rho || a.u || a.p) write_moments<L, C>(a, cell, false, rho, ua, ub, uc);
    if (COLLIDE) {
      const C inv_tau = a.tau_field ? recip(C(a.tau_field[cell])) : a.inv_tau;
      collide_for<L, T, C, HASF, SCALED>(dk, inv_tau, Fa, Fb, Fc, drho, rho, ua, ub, uc);
    }
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      // same conversion as Storage<T>::store, kept in registers for a vector store
      out[q][k] = enc_u<L, T, SCALED>(q, dk[q]);
    }
  }
  T* dst = static_cast<T*>(a.dst);
#pragma unroll
  for (int q = 0; q < L::Q; ++q)
    vstore_nt<T, V>(reinterpret_cast<T*>(reinterpret_cast<char*>(dst + q * P + seg0) + lofs),
                    out[q]);
}

// Zou-He face `which` applied in place to a post-stream state (the push
// path, where the face cells' populations only exist after the scatter).
// One thread per cell of the face.
template <class L, class T>
__global__ void __launch_bounds__(256) k_face(FwdArgs<typename Storage<T>::C> a, int which) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  const int ax = a.zh_axis[which];
  const int side = a.zh_side[which];
  const int n1 = ax == 0 ? a.nb : a.na;
  const int n2 = ax == 2 ? a.nb : a.nc;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n1 * n2) return;
  const int u = (int)(t / n2);
  const int v = (int)(t - (long long)u * n2);
  const int nax = ax == 0 ? a.na : (ax == 1 ? a.nb : a.nc);
  const int coord = side == 0 ? 0 : nax - 1;
  const int ia = ax == 0 ? coord : u;
  const int ib = ax == 0 ? u : (ax == 1 ? coord : v);
  const int ic = ax == 2 ? coord : v;
  const long long cell = ((long long)ia * a.nb + ib) * a.nc + ic;
  if (a.mask && a.mask[cell]) return;
  T* f = static_cast<T*>(a.dst);
  C d[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) d[q] = St::template load<L>(f + q * a.plane + cell, q);
  zou_he<L, C>(d, ax, side, a.zh_rho[which]);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) St::template store<L>(f + q * a.plane + cell, q, d[q]);
}

// f = feq(rho, u) (lbm.py:180-206) written as deviations.
template <class L, class T>
__global__ void __launch_bounds__(256)
    k_equilibrium(FwdArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.plane) return;
  const C rho = a.rho ? C(a.rho[cell]) : C(1);
  C ua = 0, ub = 0, uc = 0;
  if (a.u) {
    if (L::D == 3) {
      ua = C(a.u[3 * cell + 0]);
      ub = C(a.u[3 * cell + 1]);
      uc = C(a.u[3 * cell + 2]);
    } else {
      ub = C(a.u[2 * cell + 0]);
      uc = C(a.u[2 * cell + 1]);
    }
  }
  const C drho = rho - C(1);
  const C usq15 = C(1.5) * (ua * ua + ub * ub + uc * uc);
  T* dst = static_cast<T*>(a.dst);
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    const C eu = edot<L, C>(q, ua, ub, uc);
    const C dq = weight<L, C>(q) * (drho + rho * (C(3) * eu + C(4.5) * eu * eu - usq15));
    St::template store<L>(dst + q * a.plane + cell, q, dq);
  }
}

// (*G, Q) float/double <-> SoA storage.
template <class L, class T, class A, bool TO_SOA>
__global__ void __launch_bounds__(256) k_convert(FwdArgs<typename Storage<T>::C> a) {
  using C = typename Storage<T>::C;
  using St = Storage<T>;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.plane) return;
  if (TO_SOA) {
    const A* aos = static_cast<const A*>(a.src);
    T* soa = static_cast<T*>(a.dst);
#pragma unroll
    for (int q = 0; q < L::Q; ++q)
      St::template store<L>(soa + q * a.plane + cell, q,
                            C(aos[cell * L::Q + q]) - weight<L, C>(q));
  } else {
    const T* soa = static_cast<const T*>(a.src);
    A* aos = static_cast<A*>(a.dst);
#pragma unroll
    for (int q = 0; q < L::Q; ++q)
      aos[cell * L::Q + q] =
          A(St::template load<L>(soa + q * a.plane + cell, q) + weight<L, C>(q));
  }
}

// FP64 conservation sums: per-workgroup partials, then one workgroup folds
// them (deterministic order).
constexpr int kReduceBlocks = 1024;

template <class L, class T>
__global__ void __launch_bounds__(256)
    k_conserve_partial(const T* f, long long plane, double* partial) {
  using St = Storage<T>;
  double s[4] = {0, 0, 0, 0};
  for (long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x; cell < plane;
       cell += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      const double fq = double(St::template load<L>(f + q * plane + cell, q)) +
                        weight<L, double>(q);
      s[0] += fq;
      s[1] += L::ea[q] * fq;
      s[2] += L::eb[q] * fq;
      s[3] += L::ec[q] * fq;
    }
  }
  __shared__ double sh[4][256];
#pragma unroll
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = sh[k][0];
}

template <int D>
__global__ void __launch_bounds__(256) k_conserve_final(const double* partial, int nparts,
                                                        double* out) {
  __shared__ double sh[4][256];
  double s[4] = {0, 0, 0, 0};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] += partial[4 * i + k];
#pragma unroll
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0][0];
    if (D == 3) {
      out[1] = sh[1][0];
      out[2] = sh[2][0];
      out[3] = sh[3][0];
    } else {  // D2Q9 lives on axes (b, c)
      out[1] = sh[2][0];
      out[2] = sh[3][0];
      out[3] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// Byte offset of the link words in lbm_step's workspace (after one state).
inline size_t links_offset(const lbm_desc* d, const Geom& g) {
  const size_t state = (size_t)g.plane * (size_t)d->q * (size_t)storage_size(d->storage);
  return (state + 255) & ~(size_t)255;
}

template <class C>
FwdArgs<C> make_args(const lbm_desc* d, const Geom& g) {
  FwdArgs<C> a;
  std::memset(&a, 0, sizeof(a));
  a.plane = g.plane;
  a.na = g.na; a.nb = g.nb; a.nc = g.nc;
  a.bpr = g.bpr;
  a.nblocks = g.nblocks;
  a.remap = xcd_group(d, g);
  a.flat_blocks = flat_tiles(d, g);
  a.flat_staged = (a.flat_blocks > 0 && g.plane % 64 != 0 && storage_size(d->storage) <= 4) ? 1 : 0;
  if (a.flat_blocks && (d->flags & LBM_FLAG_XCD_REMAP))
    a.flat_remap = d->xcd_group > 0 ? d->xcd_group : a.flat_blocks / 8;
  a.inv_tau = C(1.0 / d->tau);
  if (d->q == 19) {
    a.Fa = C(d->force[0]); a.Fb = C(d->force[1]); a.Fc = C(d->force[2]);
    a.lid_a = C(d->lid_velocity[0]); a.lid_b = C(d->lid_velocity[1]); a.lid_c = C(d->lid_velocity[2]);
  } else {
    a.Fa = C(0); a.Fb = C(d->force[0]); a.Fc = C(d->force[1]);
    a.lid_a = C(0); a.lid_b = C(d->lid_velocity[0]); a.lid_c = C(d->lid_velocity[1]);
  }
  const int faces[2] = {d->inlet_face, d->outlet_face};
  const double rhos[2] = {d->inlet_rho, d->outlet_rho};
  a.zh_n = 0;
  for (int i = 0; i < 2; ++i) {
    if (faces[i] <= 0) continue;
    const int ref_axis = (faces[i] - 1) / 2;  // 1..6: x_min, x_max, y_min, y_max, z_min, z_max
    a.zh_axis[a.zh_n] = ref_axis + (d->q == 19 ? 0 : 1);
    a.zh_side[a.zh_n] = (faces[i] - 1) % 2;
    a.zh_rho[a.zh_n] = C(rhos[i]);
    ++a.zh_n;
  }
  return a;
}



// Vector width of the row-per-wave kernel for this lattice (0: not eligible).
// Auto (row_vec 0), from 256^3 D3Q19 measurements on MI355X: 16-bit
// storage uses the row kernel with 2 cells per lane (4-byte lane accesses:
// 5.85 TB/s against 5.27 TB/s for the one-cell-per-thread kernel); 32-bit
// storage keeps the one-cell-per-thread kernel (5.69 TB/s, the same as the
// best row variant within noise).
inline int row_width(const lbm_desc* d, const Geom& g) {
  if (d->bb_mode != LBM_BB_NONE || d->row_vec < 0) return 0;
  if (d->inlet_face > 0 || d->outlet_face > 0) return 0;  // faces live in the per-cell kernels
  if (d->row_vec == 0) {
    if (storage_size(d->storage) != 2) return 0;
    if (g.nc % 128 == 0) return 2;
    return g.nc % 64 == 0 ? 1 : 0;
  }
  return g.nc % (64 * d->row_vec) == 0 ? d->row_vec : 0;
}

template <class L, class T, int V, bool PULL, bool COLLIDE>
int launch_row(hipStream_t s, const Geom& g, const FwdArgs<typename Storage<T>::C>& a) {
  const long long waves = (long long)g.na * g.nb * (g.nc / (64 * V));
  const long long blocks = (waves + 3) / 4;
  if (blocks >= (1LL << 31)) return fail(LBM_ERR_UNSUPPORTED, "grid too large");
  // a colliding launch without a force leaves out the force terms (the
  // template argument is !COLLIDE so that a non-colliding instantiation, which
  // never takes this branch, is the same kernel as the one below)
  if (COLLIDE && !a.force_field && a.Fa == 0 && a.Fb == 0 && a.Fc == 0)
    hipLaunchKernelGGL((k_row<L, T, V, PULL, COLLIDE, !COLLIDE>), dim3((unsigned)blocks),
                       dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_row<L, T, V, PULL, COLLIDE, true>), dim3((unsigned)blocks), dim3(256),
                       0, s, a);
  return check_launch("k_row launch");
}

template <class L, class T, int BB, bool PULL, bool COLLIDE, bool STORE>
int launch_cell(hipStream_t s, const Geom& g, const FwdArgs<typename Storage<T>::C>& a) {
  if constexpr (BB == BB_NONE && STORE && !std::is_same<T, f32dev_t>::value) {
    if constexpr (sizeof(T) <= 4) {
      if (g.rowv == 4) return launch_row<L, T, 4, PULL, COLLIDE>(s, g, a);
    }
    if (g.rowv == 2) return launch_row<L, T, 2, PULL, COLLIDE>(s, g, a);
    if (g.rowv == 1) return launch_row<L, T, 1, PULL, COLLIDE>(s, g, a);
  }
  if constexpr (STORE && sizeof(T) <= 4) {
    if (a.flat_staged) {
      hipLaunchKernelGGL((k_cell_staged<L, T, BB, PULL, COLLIDE>), dim3(a.flat_blocks), dim3(256),
                         0, s, a);
      return check_launch("k_cell_staged launch");
    }
  }
  if (a.flat_blocks > 0)
    hipLaunchKernelGGL((k_cell<L, T, BB, PULL, COLLIDE, STORE>), dim3(a.flat_blocks), dim3(256), 0,
                       s, a);
  else
    hipLaunchKernelGGL((k_cell<L, T, BB, PULL, COLLIDE, STORE>), dim3(g.nblocks), dim3(g.bx), 0, s,
                       a);
  return check_launch("k_cell launch");
}

template <class L, class T, int BB>
int launch_push(hipStream_t s, const Geom& g, const FwdArgs<typename Storage<T>::C>& a) {
  hipLaunchKernelGGL((k_push<L, T, BB>), dim3(g.nblocks), dim3(g.bx), 0, s, a);
  return check_launch("k_push launch");
}

// ---------------------------------------------------------------------------
// Two steps per launch (k_pair, lbm_pair.hpp)
// ---------------------------------------------------------------------------

// Storage types the two-step launches are built for: 32-bit and 16-bit
// deviation storage (which lattices and forms use them: pair_plan,
// pair9_plan).
template <class L, class T>
constexpr bool pair_storage() {
  return std::is_same<T, float>::value || std::is_same<T, f32dev_t>::value ||
         std::is_same<T, _Float16>::value || std::is_same<T, __bf16>::value;
}
// Storage types of the one-launch update (k_pair<T, PAIR_LOCAL, PAIR_MOMENTS>).
template <class L, class T>
constexpr bool one_storage() {
  return L::Q == 19 && (std::is_same<T, float>::value || std::is_same<T, f32dev_t>::value ||
                        std::is_same<T, _Float16>::value || std::is_same<T, __bf16>::value);
}

constexpr long long kPairCUs = 256;  // compute units of one MI355X
// Rows of up to this many cells are held whole by a workgroup (at least
// three of them); longer rows are cut into tiles of kPairTile columns plus
// one halo column each side. 64 beat one step per launch by 1.16-1.39x on
// every long-row lattice measured; 128 was about even with it, 32 slower.
constexpr long long kPairWholeRows = 341;
constexpr int kPairTile = 64;

// Geometry of the D3Q19 two-step march over lattice g, or false where it
// cannot run (19 planes not addressable with 32-bit offsets). Rows of at most
// 341 cells are held whole (three or more per 1024-thread workgroup); longer
// rows are cut into 64-column tiles. Tiling rows that fit whole measured
// slower (f32dev MLUPS, whole rows against 64-column tiles: 256^3 54.5k
// against 49.1k, 192^3 50.1k against 40.0k). The chunk length minimises
// rounds of one-workgroup-per-CU launches x (chunk + 2) planes.
inline bool pair_geometry(const lbm_desc* d, const Geom& g, PairArgs* pa) {
  if (d->q != 19) return false;
  if (19LL * g.plane >= (1LL << 31)) return false;
  // whole rows up to kPairWholeRows cells, else c-tiles of kPairTile columns
  const bool tiled = g.nc > kPairWholeRows && g.nc >= kPairTile;
  const int w = tiled ? kPairTile + 2 : (int)g.nc;
  if (w > 341) return false;
  const int tc = tiled ? kPairTile : (int)g.nc;
  const long long ctiles = (g.nc + tc - 1) / tc;
  const int tb = 1024 / w - 2;
  const long long btiles = (g.nb + tb - 1) / tb;
  int la = 128;
  auto blocks = [&](int l) { return ctiles * btiles * ((g.na + l - 1) / l); };
  {
    // one workgroup per CU: a launch takes ceil(blocks / CUs) rounds of
    // (chunk + 2) planes each; take the chunk that minimises that (the
    // measured best at 256^3: 128, 256x256x160: 64, 160^3: 32, 128^3: 16)
    long long best = -1;
    for (int l = 128; l >= 16; l /= 2) {
      const long long cost = ((blocks(l) + kPairCUs - 1) / kPairCUs) * (l + 2);
      if (best < 0 || cost < best) {
        best = cost;
        la = l;
      }
    }
  }
  if (blocks(la) >= (1LL << 31)) return false;
  std::memset(pa, 0, sizeof(*pa));
  pa->na = g.na;
  pa->nb = g.nb;
  pa->nc = g.nc;
  pa->pitch = (int)g.plane;
  pa->w = w;
  pa->tc = tc;
  pa->halo = tiled ? 1 : 0;
  pa->ctiles = (int)ctiles;
  pa->tb = tb;
  pa->btiles = (int)btiles;
  pa->la = la;
  pa->nblocks = (int)blocks(la);
  // runs of 16 workgroups per XCD for whole rows (256^3, chunks of 64
  // planes: 655 us against 696 us in identity order; chunks of 128: 646 us,
  // runs of 32: 648 us). C-tiles take runs of 32, two row bands of every
  // tile column, so the halo rows between the two stay in one L2
  // (128x128x1024: reads 1.52 -> 1.41 x the state, 583 -> 580 us; runs of
  // 48-160 read 1.48-1.55 x and took 590-613 us)
  pa->remap = tiled ? 32 : 16;
  if (d->flags & LBM_FLAG_XCD_REMAP) pa->remap = d->xcd_group > 0 ? d->xcd_group : pa->nblocks / 8;
  pa->inv_tau = (float)(1.0 / d->tau);
  pa->Fa = (float)d->force[0];
  pa->Fb = (float)d->force[1];
  pa->Fc = (float)d->force[2];
  return true;
}

// The three-step march (k_tri) for 32- and 16-bit storage, periodic,
// without force or fields: c-tiles of 64 output columns with two halo
// columns each side (rows of at least 64 cells), the output array within
// the 4 GiB range of its buffer stores. The chunk length minimises rounds of
// one-workgroup-per-CU launches x (chunk + 4) planes.
inline bool tri_geometry(const lbm_desc* d, const Geom& g, PairArgs* pa) {
  if (d->q != 19 || d->bb_mode != LBM_BB_NONE) return false;
  if (d->storage == LBM_F64) return false;
  if (d->force[0] != 0.0 || d->force[1] != 0.0 || d->force[2] != 0.0) return false;
  if (g.nc < kTriTile) return false;
  if (19LL * g.plane >= (1LL << 31)) return false;
  if (20LL * g.plane * storage_size(d->storage) >= (1LL << 32)) return false;
  const long long ctiles = (g.nc + kTriTile - 1) / kTriTile;
  const long long btiles = (g.nb + kTriOut - 1) / kTriOut;
  auto blocks = [&](int l) { return ctiles * btiles * ((g.na + l - 1) / l); };
  int la = 128;
  long long best = -1;
  for (int l = 128; l >= 8; l /= 2) {
    const long long cost = ((blocks(l) + kPairCUs - 1) / kPairCUs) * (l + 4);
    if (best < 0 || cost < best) {
      best = cost;
      la = l;
    }
  }
  if (blocks(la) >= (1LL << 31)) return false;
  std::memset(pa, 0, sizeof(*pa));
  pa->na = g.na;
  pa->nb = g.nb;
  pa->nc = g.nc;
  pa->pitch = (int)g.plane;
  pa->w = kTriW;
  pa->tc = kTriTile;
  pa->halo = 2;
  pa->ctiles = (int)ctiles;
  pa->tb = kTriOut;
  pa->btiles = (int)btiles;
  pa->la = la;
  pa->nblocks = (int)blocks(la);
  pa->remap = 16;
  if (d->flags & LBM_FLAG_XCD_REMAP) pa->remap = d->xcd_group > 0 ? d->xcd_group : pa->nblocks / 8;
  pa->inv_tau = (float)(1.0 / d->tau);
  return true;
}

// Where lbm_step folds the collision and the first two fused steps into one
// launch (an even nsteps >= 4 of a sequence run two units per launch): the
// geometry above with whole 64-column tiles, unless LBM_FLAG_TRI_OFF.
// 256^3, us for the three units against (C, SC) plus a single SC launch:
// f32dev 654-658 against 480 + 455, f32 755-757 against 502 + 449, f16dev
// 534-536 against 363 + 234, bf16dev 552-553 against 347 + 233; in the
// middle of a sequence (SC, SC, SC) was no faster per unit than (SC, SC)
// (253 against 248-253 us), so it is used only there.
inline bool tri_plan(const lbm_desc* d, const Geom& g, bool fields, PairArgs* pa) {
  if (fields || (d->flags & LBM_FLAG_TRI_OFF) || g.nc % kTriTile != 0) return false;
  return tri_geometry(d, g, pa);
}

template <class T>
int launch_tri_first(hipStream_t s, const PairArgs& pa) {
  if constexpr (!std::is_same<T, double>::value) {
    hipLaunchKernelGGL((k_tri<T>), dim3((unsigned)pa.nblocks), dim3(1024), 0, s, pa);
    return check_launch("k_tri launch");
  } else {
    return fail(LBM_ERR_UNSUPPORTED, "three-unit launches are for 32- and 16-bit storage");
  }
}

// Whether the two-step march can run this call at all: D3Q19, the reference
// wall rule or none (not halfway bounce-back), and a lattice the geometry
// covers.
inline bool pair_applies(const lbm_desc* d, const Geom& g, PairArgs* pa) {
  if (d->q != 19 || d->bb_mode == LBM_BB_HALFWAY) return false;
  return pair_geometry(d, g, pa);
}

// Plan of the two-step launches of an lbm_step sequence over lattice g, or
// false where lbm_step keeps one step per launch: 32-bit storage, the march
// applies, and (without LBM_FLAG_PAIR_FORCE) the lattice is one where it
// measured faster than two fused launches: periodic with uniform force and
// tau (`fields`: the call passes a force or tau field) and no pressure faces,
// at least 2^21 cells, at least 64 cells along a and b (thinner axes
// recompute their halo planes and rows several times over), and rows either
// of 64 to 256 cells in multiples of 32 (whole rows of 288-341 cells leave a
// workgroup three rows, one of them its own output) or longer than 341 cells
// (tiled: 128x128x1024 49.1k against 36.9k MLUPS, 256x256x512 48.1k against
// 34.7k, 128x128x512 43.5k against 37.5k, 64x64x2048 44.3k against 37.3k,
// 96x96x1000 42.2k against 35.7k). Measured whole rows, f32dev, ms per step,
// two-step against the best one-step launch shape: 256^3 0.31 against 0.45,
// 512x256x128 0.338 against 0.473, 320x320x192 0.440 against 0.574, 192^3
// 0.154 against 0.203, 256x256x160 0.210 against 0.302, 160^3 0.100 against
// 0.125, 128^3 0.056 against 0.066, 64x256x128 0.056 against 0.065,
// 256x128x64 0.057 against 0.0645; below that size or with other row
// lengths it broke even or lost (200x200x100 0.117 against 0.115, 100^3
// 0.050 against 0.031, 96^3 0.045 against 0.028). With the reference wall
// rule, force or tau fields (`fields`) or pressure faces the march runs its
// general form under the same rule: on the 128x128x256 sphere channel
// (30 % walls, tau field) 0.090 against 0.120 ms per step.
//
// 16-bit storage: without LBM_FLAG_PAIR_FORCE the march runs only in its
// plain form (periodic, no force, no fields, no faces: its other forms
// spill registers with 16-bit storage), under the same shape rule. Measured
// against the row kernel (MLUPS, march / row): 256^3 f16dev 88.0k / 73.1k,
// bf16dev 89.9k / 73.4k; 128x128x1024 (tiles) f16dev 82.5k / 74.6k,
// bf16dev 76.1k / 74.3k; other shapes in the design notes.
inline bool pair_plan(const lbm_desc* d, const Geom& g, bool fields, PairArgs* pa) {
  if (d->flags & LBM_FLAG_PAIR_OFF) return false;
  if (d->storage == LBM_F64) return false;
  if (!pair_applies(d, g, pa)) return false;
  if (d->flags & LBM_FLAG_PAIR_FORCE) return true;
  // Zou-He faces: the faces form spills registers and lost to one step per
  // launch (128x128x256 sphere channel with z faces, f32: 0.149 against
  // 0.118 ms per step); a single update still takes it (0.20 against 0.37
  // ms for push + faces + moments)
  if (d->inlet_face > 0 || d->outlet_face > 0) return false;
  const bool wide = d->storage == LBM_F32 || d->storage == LBM_F32_DEV;
  if (!wide && (fields || d->bb_mode != LBM_BB_NONE || d->inlet_face > 0 ||
                d->outlet_face > 0 || d->force[0] != 0.0 || d->force[1] != 0.0 ||
                d->force[2] != 0.0))
    return false;
  if (g.nc > kPairWholeRows) {
    // 64-column tiles: only rows the tiles cover with at most 10 % of extra
    // columns (the measured ragged row, 1000 cells, computes 1024; 342 cells
    // would compute 384 and was never timed)
    const long long covered = (g.nc + kPairTile - 1) / kPairTile * kPairTile;
    if (10 * (covered - g.nc) > g.nc) return false;
  }
  return !(g.plane < (1LL << 21) || g.na < 64 || g.nb < 64 || g.nc < 64 ||
           (g.nc <= kPairWholeRows && (g.nc > 256 || g.nc % 32 != 0)));
}

// A single update (nsteps = 1) as one launch of the march, step 1 colliding
// the stored state in place of a gather and step 2 writing the streamed
// state and its moments (k_pair<T, PAIR_LOCAL, PAIR_MOMENTS>), in place of a
// push launch plus a moments pass: D3Q19, 32- or 16-bit storage, no halfway
// bounce-back, unless LBM_FLAG_PAIR_OFF.
inline bool pair_one(const lbm_desc* d, const Geom& g, PairArgs* pa) {
  if (d->flags & LBM_FLAG_PAIR_OFF) return false;
  if (d->storage == LBM_F64) return false;
  return pair_applies(d, g, pa);
}

// D2Q9 plan: periodic, 32-bit or 16-bit deviation storage (16-bit only
// under LBM_FLAG_PAIR_FORCE: 8192^2 f16dev 71.6k MLUPS against 158.8k for
// the row kernel), uniform force and tau, no pressure faces, 9 planes
// addressable with 32-bit offsets. Without LBM_FLAG_PAIR_FORCE only 32-bit
// storage and only from 2^21 cells on, where it measured faster (f32dev,
// per two steps against two one-step launches, with the march's two LDS
// halves: 8192^2 956 against 1750 us, 2048^2 82.5 against 107.8, 1536^2
// 45.1 against 68.1; 1024^2 35.6 against 39.5 but f32 even, 39.4 against
// 39.6; 512^2 24 against 14).
inline bool pair9_plan(const lbm_desc* d, const Geom& g, bool fields, Pair9Args* pa) {
  if (d->flags & LBM_FLAG_PAIR_OFF) return false;
  const bool force = (d->flags & LBM_FLAG_PAIR_FORCE) != 0;
  if (d->q != 9 || d->bb_mode != LBM_BB_NONE || fields) return false;
  if (d->storage == LBM_F64) return false;
  if (!force && d->storage != LBM_F32 && d->storage != LBM_F32_DEV) return false;
  if (d->inlet_face > 0 || d->outlet_face > 0) return false;
  if (9LL * g.plane >= (1LL << 31)) return false;
  if (!force && g.plane < (1LL << 21)) return false;
  // chunk rows by the same rounds model as the D3Q19 plan (measured best:
  // 8192^2 and 2048^2 32 rows, 4096^2 16 rows: 311 against 339 us)
  const long long ctiles = (g.nc + 959) / 960;
  auto nblk = [&](int l) { return ctiles * ((g.nb + l - 1) / l); };
  int lb = 32;
  long long best = -1;
  for (int l = 64; l >= 16; l /= 2) {
    const long long cost = ((nblk(l) + kPairCUs - 1) / kPairCUs) * (l + 2);
    if (best < 0 || cost < best) {
      best = cost;
      lb = l;
    }
  }
  const long long blocks = nblk(lb);
  if (blocks >= (1LL << 31)) return false;
  std::memset(pa, 0, sizeof(*pa));
  pa->nb = g.nb;
  pa->nc = g.nc;
  pa->pitch = (int)g.plane;
  pa->ctiles = (int)ctiles;
  pa->lb = lb;
  pa->nblocks = (int)blocks;
  pa->inv_tau = (float)(1.0 / d->tau);
  pa->Fb = (float)d->force[0];
  pa->Fc = (float)d->force[1];
  return true;
}

// One two-step launch of kind (S1, S2) in the form the call needs.
template <class T, int S1, int S2>
int launch_pair_gen(hipStream_t s, const PairArgs& pa) {
  const dim3 grid((unsigned)pa.nblocks), block(1024);
  switch (pair_form(pa)) {
    case PAIR_PLAIN:
      // buffer stores when every population and a plane past the last end
      // within 4 GiB of the array's base
      if (20LL * pa.pitch * (long long)sizeof(T) < (1LL << 32))
        hipLaunchKernelGGL((k_pair<T, S1, S2, PAIR_PLAIN, true>), grid, block, 0, s, pa);
      else
        hipLaunchKernelGGL((k_pair<T, S1, S2, PAIR_PLAIN>), grid, block, 0, s, pa);
      break;
    case PAIR_FORCE:
      hipLaunchKernelGGL((k_pair<T, S1, S2, PAIR_FORCE>), grid, block, 0, s, pa);
      break;
    case PAIR_GEN:
      if (20LL * pa.pitch * (long long)sizeof(T) < (1LL << 32))
        hipLaunchKernelGGL((k_pair<T, S1, S2, PAIR_GEN, true>), grid, block, 0, s, pa);
      else
        hipLaunchKernelGGL((k_pair<T, S1, S2, PAIR_GEN>), grid, block, 0, s, pa);
      break;
    default:
      hipLaunchKernelGGL((k_pair<T, S1, S2, PAIR_GEN_FACES>), grid, block, 0, s, pa);
  }
  return check_launch("k_pair launch");
}

// The two-step launch of each lattice: its argument block, plan and kernel.
template <class L> struct PairOf;
template <> struct PairOf<Lattice<19>> {
  using Args = PairArgs;
  static bool plan(const lbm_desc* d, const Geom& g, bool f, Args* a) { return pair_plan(d, g, f, a); }
  template <class T>
  static int launch(hipStream_t s, const Args& pa) {
    return launch_pair_gen<T, PAIR_GATHER, PAIR_COLLIDE>(s, pa);
  }
};
template <> struct PairOf<Lattice<9>> {
  using Args = Pair9Args;
  static bool plan(const lbm_desc* d, const Geom& g, bool f, Args* a) { return pair9_plan(d, g, f, a); }
  template <class T>
  static int launch(hipStream_t s, const Args& pa) {
    hipLaunchKernelGGL((k_pair9<T>), dim3((unsigned)pa.nblocks), dim3(1024), 0, s, pa);
    return check_launch("k_pair9 launch");
  }
};

// The general fields of a two-step launch from the single-step arguments of
// the same call: mask (reference wall rule only), force and tau fields,
// faces.
template <class C>
void pair_general(PairArgs* pa, const FwdArgs<C>& a, bool bb_ref) {
  pa->mask = bb_ref ? a.mask : nullptr;
  pa->force_field = a.force_field;
  pa->tau_field = a.tau_field;
  pa->zh_n = a.zh_n;
  for (int i = 0; i < 2; ++i) {
    pa->zh_axis[i] = a.zh_axis[i];
    pa->zh_side[i] = a.zh_side[i];
    pa->zh_rho[i] = float(a.zh_rho[i]);
  }
}

// Launches of a pull sequence of nsteps steps. Its units are one collision
// C, then nsteps - 1 fused steps SC, then one stream + moments S.
//  * One step per launch: C, SC x (nsteps - 1), S.
//  * D2Q9 two-step launches: C, pairs of SC, an odd SC, S.
//  * D3Q19 two-step launches (`fold`): every launch but at most one is a
//    two-step launch. nsteps = 1: (C, S) in one launch; nsteps = 2: C, then
//    (SC, S); nsteps >= 3: (C, SC), the odd SC when nsteps is even, pairs of
//    SC, (SC, S).
//  * With `tri` (an even nsteps >= 4 where tri_plan applies): (C, SC, SC),
//    pairs of SC, (SC, S), every launch a multi-unit one.
struct PullSchedule {
  int launches;   // all launches
  int pairs;      // launches of two or three units among them
  int npair;      // (SC, SC) launches
  int nsingle;    // single SC launches
  bool fold;      // the first and last units run inside two-step launches
  bool tri;       // the first launch is (C, SC, SC)
};
inline PullSchedule pull_schedule(int nsteps, bool pair, bool fold, bool tri = false) {
  PullSchedule sc;
  sc.fold = pair && fold;
  sc.tri = false;
  if (sc.fold && tri && nsteps >= 4 && nsteps % 2 == 0) {
    sc.tri = true;
    sc.nsingle = 0;
    sc.npair = (nsteps - 4) / 2;
    sc.launches = sc.pairs = 2 + sc.npair;
    return sc;
  }
  if (!sc.fold) {
    const int fused = nsteps - 1;
    sc.npair = pair ? fused / 2 : 0;
    sc.nsingle = fused - 2 * sc.npair;
    sc.launches = 2 + sc.npair + sc.nsingle;
    sc.pairs = sc.npair;
    return sc;
  }
  if (nsteps == 1) {
    sc.npair = sc.nsingle = 0;
    sc.launches = sc.pairs = 1;
  } else if (nsteps == 2) {
    sc.npair = sc.nsingle = 0;  // the lone collision runs single, then (SC, S)
    sc.launches = 2;
    sc.pairs = 1;
  } else {
    sc.nsingle = nsteps % 2 == 0 ? 1 : 0;
    sc.npair = (nsteps - 3 - sc.nsingle) / 2;
    sc.launches = 2 + sc.nsingle + sc.npair;
    sc.pairs = 2 + sc.npair;
  }
  return sc;
}

// Pull sequence (launch j writes buf[(launches - 1 - j) & 1], so the last
// writes f_out and reads `work`; a one-launch sequence needs no workspace):
// collide (f_in -> B0), fused pull+collide (B -> B'), final pull + moments
// (B -> f_out), with two fused steps per launch where `pair` is given and,
// for D3Q19, the first and last units folded into two-step launches.
template <class L, class T, int BB>
int step_pull(hipStream_t s, const Geom& g, FwdArgs<typename Storage<T>::C> a,
              const void* f_in, void* f_out, void* work, int nsteps,
              float* rho, float* u, float* p,
              const typename PairOf<L>::Args* pair = nullptr, const PairArgs* tri = nullptr) {
  using C = typename Storage<T>::C;
  int rc;
  constexpr bool kFold = L::Q == 19 && (BB == BB_NONE || BB == BB_REF);
  const PullSchedule sc = pull_schedule(nsteps, pair != nullptr, kFold, tri != nullptr);
  void* buf[2] = {f_out, work};
  int j = 0;
  auto next = [&]() { return buf[(sc.launches - 1 - j++) & 1]; };
  FwdArgs<C> k = a;
  k.rho = k.u = k.p = nullptr;
  if constexpr (L::Q == 19 && (BB == BB_NONE || BB == BB_REF) && one_storage<L, T>()) {
    if (sc.fold) {
      PairArgs pa = *pair;
      pair_general(&pa, a, BB == BB_REF);
      if (sc.launches == 1) {
        // (C, S): the whole update in one launch
        pa.src = f_in;
        pa.dst = next();
        pa.rho = rho; pa.u = u; pa.p = p;
        return launch_pair_gen<T, PAIR_LOCAL, PAIR_MOMENTS>(s, pa);
      }
      if constexpr (pair_storage<L, T>()) {
        const void* cur = f_in;
        if (nsteps == 2) {
          k.src = f_in;
          k.dst = next();
          if ((rc = launch_cell<L, T, BB, false, true, true>(s, g, k))) return rc;
          cur = k.dst;
        } else if (sc.tri) {
          PairArgs ta = *tri;
          ta.src = f_in;
          ta.dst = next();
          if ((rc = launch_tri_first<T>(s, ta))) return rc;
          cur = ta.dst;
          for (int i = 0; i < sc.npair; ++i) {
            pa.src = cur;
            pa.dst = next();
            rc = launch_pair_gen<T, PAIR_GATHER, PAIR_COLLIDE>(s, pa);
            if (rc) return rc;
            cur = pa.dst;
          }
        } else {
          pa.src = f_in;
          pa.dst = next();
          if ((rc = launch_pair_gen<T, PAIR_LOCAL, PAIR_COLLIDE>(s, pa))) return rc;
          cur = pa.dst;
          for (int i = 0; i < sc.nsingle; ++i) {
            k.src = cur;
            k.dst = next();
            if ((rc = launch_cell<L, T, BB, true, true, true>(s, g, k))) return rc;
            cur = k.dst;
          }
          for (int i = 0; i < sc.npair; ++i) {
            pa.src = cur;
            pa.dst = next();
            rc = launch_pair_gen<T, PAIR_GATHER, PAIR_COLLIDE>(s, pa);
            if (rc) return rc;
            cur = pa.dst;
          }
        }
        pa.src = cur;
        pa.dst = next();
        pa.rho = rho; pa.u = u; pa.p = p;
        return launch_pair_gen<T, PAIR_GATHER, PAIR_MOMENTS>(s, pa);
      }
    }
  }
  k.src = f_in;
  k.dst = next();
  if ((rc = launch_cell<L, T, BB, false, true, true>(s, g, k))) return rc;
  if constexpr (BB == BB_NONE && pair_storage<L, T>()) {
    if (sc.npair) {
      typename PairOf<L>::Args pa = *pair;
      for (int i = 0; i < sc.npair; ++i) {
        pa.src = k.dst;
        pa.dst = next();
        if ((rc = PairOf<L>::template launch<T>(s, pa))) return rc;
        k.dst = pa.dst;
      }
    }
  }
  for (int i = 0; i < sc.nsingle; ++i) {
    k.src = k.dst;
    k.dst = next();
    if ((rc = launch_cell<L, T, BB, true, true, true>(s, g, k))) return rc;
  }
  k.src = work;
  k.dst = f_out;
  k.rho = rho; k.u = u; k.p = p;
  return launch_cell<L, T, BB, true, false, true>(s, g, k);
}

template <class L, class T, int BB>
int step_push(hipStream_t s, const Geom& g, FwdArgs<typename Storage<T>::C> a,
              const void* f_in, void* f_out, void* work, int nsteps,
              float* rho, float* u, float* p) {
  using C = typename Storage<T>::C;
  int rc;
  void* buf[2] = {f_out, work};  // buf[(nsteps-1-i) & 1]... state after step i+1 -> buf[(nsteps-1-i)&1]
  FwdArgs<C> k = a;
  k.rho = k.u = k.p = nullptr;
  const void* cur = f_in;
  for (int i = 0; i < nsteps; ++i) {
    k.src = cur;
    k.dst = buf[(nsteps - 1 - i) & 1];
    if ((rc = launch_push<L, T, BB>(s, g, k))) return rc;
    for (int face = 0; face < k.zh_n; ++face) {
      const int ax = k.zh_axis[face];
      const long long cells = (long long)(ax == 0 ? g.nb : g.na) * (ax == 2 ? g.nb : g.nc);
      hipLaunchKernelGGL((k_face<L, T>), dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, k,
                         face);
      if ((rc = check_launch("k_face launch"))) return rc;
    }
    cur = k.dst;
  }
  if (rho || u || p) {
    k.src = f_out;
    k.dst = nullptr;
    k.rho = rho; k.u = u; k.p = p;
    return launch_cell<L, T, BB_NONE, false, false, false>(s, g, k);
  }
  return LBM_OK;
}

// Halfway link words (k_links) for the lattice g, two words per cell.
template <class L, class T>
struct LinksOp {
  static int run(hipStream_t s, const Geom& g, const uint8_t* mask, uint32_t* links) {
    hipLaunchKernelGGL((k_links<L>), dim3(g.nblocks), dim3(g.bx), 0, s, mask, links,
                       links + g.plane, g.na, g.nb, g.nc, g.bpr);
    return check_launch("k_links launch");
  }
};

template <class L, class T>
int step_typed(hipStream_t s, const lbm_desc* d, const Geom& g, const void* f_in,
               void* f_out, void* work, const uint8_t* mask, const float* ff,
               const float* tf, float* rho, float* u, float* p, int nsteps,
               int scheme) {
  using C = typename Storage<T>::C;
  FwdArgs<C> a = make_args<C>(d, g);
  a.mask = d->bb_mode == LBM_BB_NONE ? nullptr : mask;
  a.force_field = ff;
  a.tau_field = tf;
  Geom gr = g;
  gr.rowv = scheme == LBM_SCHEME_PULL ? row_width(d, g) : 0;
  if (scheme == LBM_SCHEME_PULL) {
    // two-step launches: a single update as one launch (pair_one), longer
    // sequences where pair_plan measured them faster (or PAIR_FORCE)
    typename PairOf<L>::Args pa;
    PairArgs ta;
    bool pair = false, tri = false;
    if constexpr (L::Q == 19) {
      if (nsteps == 1) pair = one_storage<L, T>() && pair_one(d, g, &pa);
      else pair = pair_storage<L, T>() && pair_plan(d, g, ff || tf, &pa);
      tri = pair && nsteps >= 4 && !std::is_same<T, double>::value && tri_plan(d, g, ff || tf, &ta);
    } else {
      pair = pair_storage<L, T>() && PairOf<L>::plan(d, g, ff || tf, &pa);
    }
    switch (d->bb_mode) {
      case LBM_BB_NONE:
        return step_pull<L, T, BB_NONE>(s, gr, a, f_in, f_out, work, nsteps, rho, u, p,
                                        pair ? &pa : nullptr, tri ? &ta : nullptr);
      case LBM_BB_REFERENCE:
        return step_pull<L, T, BB_REF>(s, g, a, f_in, f_out, work, nsteps, rho, u, p,
                                       pair ? &pa : nullptr);
      default: {
        uint32_t* links = reinterpret_cast<uint32_t*>(static_cast<char*>(work) + links_offset(d, g));
        if (int rc = LinksOp<L, T>::run(s, g, mask, links)) return rc;
        a.links = links;
        a.links_moving = links + g.plane;
        return step_pull<L, T, BB_HALF_LINKS>(s, g, a, f_in, f_out, work, nsteps, rho, u, p);
      }
    }
  }
  switch (d->bb_mode) {
    case LBM_BB_NONE: return step_push<L, T, BB_NONE>(s, g, a, f_in, f_out, work, nsteps, rho, u, p);
    case LBM_BB_REFERENCE: return step_push<L, T, BB_REF>(s, g, a, f_in, f_out, work, nsteps, rho, u, p);
    default: return step_push<L, T, BB_HALF>(s, g, a, f_in, f_out, work, nsteps, rho, u, p);
  }
}

template <template <class, class> class F, class... Args>
int dispatch(const lbm_desc* d, Args&&... args) {
  if (d->q == 19) {
    switch (d->storage) {
      case LBM_F32: return F<Lattice<19>, float>::run(args...);
      case LBM_F64: return F<Lattice<19>, double>::run(args...);
      case LBM_F16_DEV: return F<Lattice<19>, _Float16>::run(args...);
      case LBM_F32_DEV: return F<Lattice<19>, f32dev_t>::run(args...);
      default: return F<Lattice<19>, __bf16>::run(args...);
    }
  }
  switch (d->storage) {
    case LBM_F32: return F<Lattice<9>, float>::run(args...);
    case LBM_F64: return F<Lattice<9>, double>::run(args...);
    case LBM_F16_DEV: return F<Lattice<9>, _Float16>::run(args...);
    case LBM_F32_DEV: return F<Lattice<9>, f32dev_t>::run(args...);
    default: return F<Lattice<9>, __bf16>::run(args...);
  }
}

template <class L, class T>
struct StepOp {
  static int run(hipStream_t s, const lbm_desc* d, const Geom& g, const void* f_in,
                 void* f_out, void* work, const uint8_t* mask, const float* ff,
                 const float* tf, float* rho, float* u, float* p, int nsteps, int scheme) {
    return step_typed<L, T>(s, d, g, f_in, f_out, work, mask, ff, tf, rho, u, p, nsteps, scheme);
  }
};

template <class L, class T, int BB>
int pull_phase_bb(hipStream_t s, const Geom& g, const FwdArgs<typename Storage<T>::C>& a,
                  int phase) {
  if (phase == 0) return launch_cell<L, T, BB, false, true, true>(s, g, a);
  if (phase == 1) return launch_cell<L, T, BB, true, true, true>(s, g, a);
  return launch_cell<L, T, BB, true, false, true>(s, g, a);
}

template <class L, class T>
struct PullPhaseOp {
  static int run(hipStream_t s, const lbm_desc* d, const Geom& g, int phase, const void* src,
                 void* dst, const uint8_t* mask, const uint32_t* links, const float* ff,
                 const float* tf, float* rho, float* u, float* p) {
    using C = typename Storage<T>::C;
    FwdArgs<C> a = make_args<C>(d, g);
    a.src = src;
    a.dst = dst;
    a.mask = d->bb_mode == LBM_BB_NONE ? nullptr : mask;
    a.links = links;
    a.links_moving = links ? links + g.plane : nullptr;
    a.force_field = ff;
    a.tau_field = tf;
    if (phase == 2) {
      a.rho = rho; a.u = u; a.p = p;
    }
    if (phase == 3 && L::Q == 9) {
      typename PairOf<L>::Args pa;
      if (!pair_storage<L, T>() || !PairOf<L>::plan(d, g, ff || tf, &pa))
        return fail(LBM_ERR_UNSUPPORTED,
                    "phase 3 (two steps per launch) does not apply to this lattice");
      if constexpr (pair_storage<L, T>()) {
        pa.src = src;
        pa.dst = dst;
        return PairOf<L>::template launch<T>(s, pa);
      }
    }
    if (phase == 7) {
      // D3Q19 three-unit launch (C, SC, SC)
      PairArgs ta;
      if (L::Q == 19 && !ff && !tf && tri_geometry(d, g, &ta)) {
        ta.src = src;
        ta.dst = dst;
        return launch_tri_first<T>(s, ta);
      }
      return fail(LBM_ERR_UNSUPPORTED,
                  "phase 7 (three units per launch) does not apply to this call");
    }
    if (phase >= 3) {
      // D3Q19 two-step launches: 3 (SC, SC), 4 (C, SC), 5 (SC, S), 6 (C, S)
      PairArgs pa;
      const bool ok = phase == 6 ? one_storage<L, T>() : pair_storage<L, T>();
      if (!ok || !pair_applies(d, g, &pa))
        return fail(LBM_ERR_UNSUPPORTED, "phase " + std::to_string(phase) +
                                             " (two steps per launch) does not apply to this lattice");
      if constexpr (one_storage<L, T>()) {
        pair_general(&pa, a, d->bb_mode == LBM_BB_REFERENCE);
        pa.src = src;
        pa.dst = dst;
        if (phase >= 5) {
          pa.rho = rho; pa.u = u; pa.p = p;
        }
        if (phase == 6) return launch_pair_gen<T, PAIR_LOCAL, PAIR_MOMENTS>(s, pa);
        if constexpr (pair_storage<L, T>()) {
          if (phase == 3) return launch_pair_gen<T, PAIR_GATHER, PAIR_COLLIDE>(s, pa);
          if (phase == 4) return launch_pair_gen<T, PAIR_LOCAL, PAIR_COLLIDE>(s, pa);
          return launch_pair_gen<T, PAIR_GATHER, PAIR_MOMENTS>(s, pa);
        }
      }
      return fail(LBM_ERR_UNSUPPORTED, "two steps per launch do not apply to this lattice");
    }
    Geom gr = g;
    gr.rowv = row_width(d, g);
    switch (d->bb_mode) {
      case LBM_BB_NONE: return pull_phase_bb<L, T, BB_NONE>(s, gr, a, phase);
      case LBM_BB_REFERENCE: return pull_phase_bb<L, T, BB_REF>(s, gr, a, phase);
      default:
        if (links) return pull_phase_bb<L, T, BB_HALF_LINKS>(s, gr, a, phase);
        return pull_phase_bb<L, T, BB_HALF>(s, gr, a, phase);
    }
  }
};

template <class L, class T>
struct MomentsOp {
  static int run(hipStream_t s, const lbm_desc* d, const Geom& g, const void* f,
                 const uint8_t* mask, const float* ff, float* rho, float* u, float* p) {
    using C = typename Storage<T>::C;
    FwdArgs<C> a = make_args<C>(d, g);
    a.src = f;
    a.mask = mask;
    a.force_field = ff;
    a.rho = rho; a.u = u; a.p = p;
    return launch_cell<L, T, BB_NONE, false, false, false>(s, g, a);
  }
};

template <class L, class T>
struct EquilibriumOp {
  static int run(hipStream_t s, const lbm_desc* d, const Geom& g, const float* rho,
                 const float* u, void* f_out) {
    using C = typename Storage<T>::C;
    FwdArgs<C> a = make_args<C>(d, g);
    a.rho = const_cast<float*>(rho);
    a.u = const_cast<float*>(u);
    a.dst = f_out;
    const int nb = (int)((g.plane + 255) / 256);
    hipLaunchKernelGGL((k_equilibrium<L, T>), dim3(nb), dim3(256), 0, s, a);
    return check_launch("k_equilibrium launch");
  }
};

template <class L, class T>
struct ConvertOp {
  static int run(hipStream_t s, const lbm_desc* d, const Geom& g, int to_soa, int f64,
                 const void* src, void* dst) {
    using C = typename Storage<T>::C;
    FwdArgs<C> a = make_args<C>(d, g);
    a.src = src;
    a.dst = dst;
    const int nb = (int)((g.plane + 255) / 256);
    if (to_soa && f64) hipLaunchKernelGGL((k_convert<L, T, double, true>), dim3(nb), dim3(256), 0, s, a);
    else if (to_soa) hipLaunchKernelGGL((k_convert<L, T, float, true>), dim3(nb), dim3(256), 0, s, a);
    else if (f64) hipLaunchKernelGGL((k_convert<L, T, double, false>), dim3(nb), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_convert<L, T, float, false>), dim3(nb), dim3(256), 0, s, a);
    return check_launch("k_convert launch");
  }
};

template <class L, class T>
struct ConserveOp {
  static int run(hipStream_t s, const lbm_desc*, const Geom& g, const void* f, void* work,
                 double* out) {
    double* partial = static_cast<double*>(work);
    long long nb = (g.plane + 255) / 256;
    const int nparts = (int)(nb < kReduceBlocks ? nb : kReduceBlocks);
    hipLaunchKernelGGL((k_conserve_partial<L, T>), dim3(nparts), dim3(256), 0, s,
                       static_cast<const T*>(f), g.plane, partial);
    int rc = check_launch("k_conserve_partial launch");
    if (rc) return rc;
    hipLaunchKernelGGL((k_conserve_final<L::D>), dim3(1), dim3(256), 0, s,
                       (const double*)partial, nparts, out);
    return check_launch("k_conserve_final launch");
  }
};

}  // namespace mlbm

using namespace mlbm;

// Partial sums of lbm_conservation_static: 4 doubles per workgroup of
// k_conserve_partial, at most kReduceBlocks workgroups (32 KiB). One copy per
// GPU, alive until the process ends. Host code must not pass `scratch`
// itself to a kernel (on the host it names a stand-in, not device memory):
// it takes the device address from scratch_address().
static __device__ __align__(256) double scratch[4 * kReduceBlocks];
constexpr size_t kScratchBytes = sizeof(double) * 4 * kReduceBlocks;

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress) and kept.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(LBM_ERR_UNSUPPORTED, "device ordinal " + std::to_string(dev) + " above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return LBM_OK;
}
}  // namespace

extern "C" {

int lbm_version(void) { return LBM_API_VERSION; }

const char* lbm_last_error(void) { return g_last_error.c_str(); }

size_t lbm_state_bytes(const lbm_desc* d) {
  Geom g;
  if (check_desc(d, &g)) return 0;
  return (size_t)g.plane * (size_t)d->q * (size_t)storage_size(d->storage);
}

// The scheme and launches of an lbm_step call. Auto: the pull sequence,
// measured 15-20 % faster per step than push, and for a single D3Q19 step
// its one-launch form (pair_one); push for a single step elsewhere (the
// pull sequence needs two launches, push one plus the moments pass).
struct CallPlan {
  int scheme;
  bool pair;
  PullSchedule sc;
};
static CallPlan call_plan(const lbm_desc* d, const Geom& g, int nsteps, bool fields) {
  CallPlan cp;
  PairArgs pa;
  Pair9Args pa9;
  const bool one = d->q == 19 && nsteps == 1 && pair_one(d, g, &pa);
  cp.scheme = d->scheme;
  if (cp.scheme == LBM_SCHEME_AUTO)
    cp.scheme = (nsteps <= 1 && !one) ? LBM_SCHEME_PUSH : LBM_SCHEME_PULL;
  cp.pair = false;
  bool tri = false;
  if (cp.scheme == LBM_SCHEME_PULL) {
    if (d->q == 19) cp.pair = nsteps == 1 ? one : pair_plan(d, g, fields, &pa);
    else cp.pair = pair9_plan(d, g, fields, &pa9);
    PairArgs ta;
    tri = d->q == 19 && cp.pair && nsteps >= 4 && tri_plan(d, g, fields, &ta);
  }
  cp.sc = pull_schedule(nsteps, cp.pair, d->q == 19 && d->bb_mode != LBM_BB_HALFWAY, tri);
  return cp;
}

size_t lbm_workspace_size(const lbm_desc* d, int nsteps) {
  Geom g;
  if (check_desc(d, &g)) return 0;
  const CallPlan cp = call_plan(d, g, nsteps < 1 ? 1 : nsteps, false);
  if (cp.scheme == LBM_SCHEME_PUSH && nsteps <= 1) return 0;
  if (cp.scheme == LBM_SCHEME_PULL && d->bb_mode == LBM_BB_HALFWAY)
    return links_offset(d, g) + 2 * (size_t)g.plane * sizeof(uint32_t);
  if (cp.scheme == LBM_SCHEME_PULL && cp.sc.launches == 1) return 0;
  return (size_t)g.plane * (size_t)d->q * (size_t)storage_size(d->storage);
}

int lbm_step(void* stream, const lbm_desc* d, const void* f_in, void* f_out, void* work,
             const uint8_t* mask, const float* force_field, const float* tau_field, float* rho,
             float* u, float* p, int nsteps) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (nsteps < 1) return fail(LBM_ERR_INVALID, "nsteps must be >= 1");
  if (!f_in || !f_out) return fail(LBM_ERR_INVALID, "f_in and f_out must be non-NULL");
  if (f_in == f_out) return fail(LBM_ERR_INVALID, "f_in and f_out must be distinct buffers");
  if (d->bb_mode != LBM_BB_NONE && !mask)
    return fail(LBM_ERR_INVALID, "bb_mode needs a wall mask");
  const int scheme = call_plan(d, g, nsteps, force_field || tau_field).scheme;
  if (lbm_workspace_size(d, nsteps) > 0 && !work)
    return fail(LBM_ERR_INVALID, "this step needs a workspace of lbm_workspace_size() bytes");
  if (work && (work == f_in || work == f_out))
    return fail(LBM_ERR_INVALID, "work must not alias f_in or f_out");
  return dispatch<StepOp>(d, (hipStream_t)stream, d, g, f_in, f_out, work, mask, force_field,
                          tau_field, rho, u, p, nsteps, scheme);
}

int lbm_step_schedule(const lbm_desc* d, int nsteps, int with_fields, int* launches,
                      int* pair_launches) {
  g_last_error.clear();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (nsteps < 1) return fail(LBM_ERR_INVALID, "nsteps must be >= 1");
  int nl, np = 0;
  const CallPlan cp = call_plan(d, g, nsteps, with_fields != 0);
  if (cp.scheme == LBM_SCHEME_PUSH) {
    nl = nsteps + 1;  // one launch per step, then the moments pass
  } else {
    nl = cp.sc.launches;
    np = cp.sc.pairs;
  }
  if (launches) *launches = nl;
  if (pair_launches) *pair_launches = np;
  return LBM_OK;
}

int lbm_pull_phase(void* stream, const lbm_desc* d, int phase, const void* src, void* dst,
                   const uint8_t* mask, const uint32_t* links, const float* force_field,
                   const float* tau_field, float* rho, float* u, float* p) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (phase < 0 || phase > 7) return fail(LBM_ERR_INVALID, "phase must be 0 to 7");
  if (!src || !dst) return fail(LBM_ERR_INVALID, "src and dst must be non-NULL");
  if (src == dst) return fail(LBM_ERR_INVALID, "src and dst must be distinct buffers");
  if (d->bb_mode != LBM_BB_NONE && !mask)
    return fail(LBM_ERR_INVALID, "bb_mode needs a wall mask");
  if (links && d->bb_mode != LBM_BB_HALFWAY)
    return fail(LBM_ERR_INVALID, "links are for LBM_BB_HALFWAY only");
  return dispatch<PullPhaseOp>(d, (hipStream_t)stream, d, g, phase, src, dst, mask, links,
                               force_field, tau_field, rho, u, p);
}

size_t lbm_links_size(const lbm_desc* d) {
  Geom g;
  if (check_desc(d, &g)) return 0;
  return 2 * (size_t)g.plane * sizeof(uint32_t);
}

int lbm_links_build(void* stream, const lbm_desc* d, const uint8_t* mask, uint32_t* links) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (!mask || !links) return fail(LBM_ERR_INVALID, "mask and links must be non-NULL");
  if (d->q == 19) return LinksOp<Lattice<19>, float>::run((hipStream_t)stream, g, mask, links);
  return LinksOp<Lattice<9>, float>::run((hipStream_t)stream, g, mask, links);
}

int lbm_moments(void* stream, const lbm_desc* d, const void* f, const uint8_t* mask,
                const float* force_field, float* rho, float* u, float* p) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (!f) return fail(LBM_ERR_INVALID, "f must be non-NULL");
  return dispatch<MomentsOp>(d, (hipStream_t)stream, d, g, f, mask, force_field, rho, u, p);
}

int lbm_equilibrium(void* stream, const lbm_desc* d, const float* rho, const float* u,
                    void* f_out) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (!f_out) return fail(LBM_ERR_INVALID, "f_out must be non-NULL");
  return dispatch<EquilibriumOp>(d, (hipStream_t)stream, d, g, rho, u, f_out);
}

int lbm_convert(void* stream, const lbm_desc* d, int to_soa, int aos_is_f64, const void* src,
                void* dst) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (!src || !dst) return fail(LBM_ERR_INVALID, "src and dst must be non-NULL");
  if (src == dst) return fail(LBM_ERR_INVALID, "src and dst must be distinct buffers");
  return dispatch<ConvertOp>(d, (hipStream_t)stream, d, g, to_soa, aos_is_f64, src, dst);
}

size_t lbm_reduce_workspace_size(const lbm_desc*) {
  return sizeof(double) * 4 * kReduceBlocks;
}

int lbm_conservation_static(void* stream, const lbm_desc* d, const void* f, double* out) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (!f || !out) return fail(LBM_ERR_INVALID, "f and out must be non-NULL");
  if (lbm_reduce_workspace_size(d) > kScratchBytes)
    return fail(LBM_ERR_UNSUPPORTED, "the reduction needs more than the 32 KiB static scratch");
  void* work = nullptr;
  if ((rc = scratch_address(&work))) return rc;
  return dispatch<ConserveOp>(d, (hipStream_t)stream, d, g, f, work, out);
}

int lbm_conservation(void* stream, const lbm_desc* d, const void* f, void* work, double* out) {
  g_last_error.clear();
  (void)hipGetLastError();
  Geom g;
  int rc = check_desc(d, &g);
  if (rc) return rc;
  if (!f || !work || !out) return fail(LBM_ERR_INVALID, "f, work and out must be non-NULL");
  return dispatch<ConserveOp>(d, (hipStream_t)stream, d, g, f, work, out);
}

}  // extern "C"
