// Three units of the pull sequence in one launch for periodic D3Q19
// lattices without a force (temporal blocking through LDS, one level deeper
// than lbm_pair.hpp): the collision C of a stored state f_n and two fused
// steps SC, giving the post-collision state C(f_{n+2}). Included by
// lbm_forward.hip.
//
// The two-step march moves the state across HBM once per two units and,
// in 32-bit storage, is bound by that traffic: with its collisions taken out
// it runs within 3 % of its time. Three units per pass move it once per
// three, and pay for it in arithmetic: 3.67 collisions per output cell
// against the two-step march's 3.0 at 256^3. In the middle of a sequence,
// as (SC, SC, SC), that came out even per unit (256^3 f32dev: 253 us against
// 248-253 for the two-step launch; with the traffic gone the march is bound
// by issue, two barriers per plane), so lbm_step uses k_tri only where it
// replaces a single-unit launch: the first launch of an even sequence,
// (C, SC, SC), 654-658 us against 480 + 455 us for (C, SC) and a single SC
// (f32dev; every 32- and 16-bit storage gains, the figures at tri_plan).
//
// A 1024-thread workgroup owns 15 rows of a c-tile: 64 output columns plus
// two halo columns each side (kTriW = 68 cells), and marches along axis a
// over a chunk of planes. Each plane p runs
//
//   step 1 at plane p, every cell: load f_n at the cell (the next plane's
//          loads are issued first, so they overlap this plane's
//          arithmetic), collide;
//   step 2 at plane p - 1 on rows 1 .. 13 and columns 1 .. 66: gather step
//          1's results through exchange buffer x1, collide;
//   step 3 at plane p - 2 on rows 2 .. 12 and the 64 inner columns: gather
//          step 2's results through x2, collide, store.
//
// As in the two-step march, each thread keeps its own cell's results of the
// last two planes of steps 1 and 2 in registers and writes, per plane, the
// e_a = -1 values of the newest plane, the e_a = 0 values of the one before
// and the e_a = +1 values of the one before that. x1 and x2 are single
// buffers of 64 KB each (two double buffers would not fit in LDS), so a
// plane takes two barriers: after the x1 writes (step 2 may read) and after
// the x2 writes (step 3 may read, and every step-2 read of x1 is done, so the
// next plane may write x1). The next plane writes x2 after its first
// barrier, which every wave reaches after this plane's step 3.
//
// Threads 0 .. 959 are the inner columns, one row per wave (aligned 64-cell
// lines); wave 15 holds the 60 halo cells. Each row wave runs with a fixed
// role (steps 1-3, 1-2 or 1 only), so no store sits in a branch and the
// wait for the next plane's loads never waits for stores. Wave 15 has a role
// of its own (step 1, step 2 per lane, no step 3 and no store), which took
// the f32dev launch from 695-699 to 654-658 us and keeps float32 storage
// from spilling in the loop (bfloat16's halo wave still spills a few
// registers there); the waves of a c-tile cut by the lattice's edge decide
// per lane.
//
// Every update is the same inlined moments + collision as k_cell's, and the
// results of steps 1 and 2 pass through the storage type's encode/decode as
// a store and reload would, so the output equals three k_cell launches bit
// for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lbm_pair.hpp"

namespace mlbm {

constexpr int kTriTile = 64;            // output columns per row
constexpr int kTriW = kTriTile + 4;     // cells per row
constexpr int kTriRows = 1024 / kTriW;  // 15
constexpr int kTriOut = kTriRows - 4;   // 11 output rows

// Roles: TRI_MIXED decides per lane; the others hold for every lane of a wave.
// TRI_H is the wave of halo columns (lanes 960-1023) of a tile the lattice
// does not cut: step 1 on every lane, step 2 per lane, never step 3 or a
// store.
enum { TRI_MIXED = 0, TRI_S1 = 1, TRI_S2 = 2, TRI_S3 = 3, TRI_H = 4 };

template <class T, int ROLE>
__device__ __forceinline__ void tri_march(const PairArgs& a, int ct, int bt, int chunk,
                                          float (&x1)[16][1024], float (&x2)[16][1024]) {
  using L = Lattice<19>;
  using C = float;
  static_assert(sizeof(typename Storage<T>::C) == 4, "float compute only");
  constexpr bool SCALED = ScaleOf<T>::v != 0;
  constexpr int LS = SCALED ? ScaleOf<T>::v : 0;
  constexpr int W = kTriW;
  const int a0 = chunk * a.la;
  const int a1 = a0 + a.la < a.na ? a0 + a.la : a.na;  // output planes [a0, a1)
  const int b0 = bt * kTriOut;
  const int nc = a.nc, na = a.na, nb = a.nb, P = a.pitch;
  const int tid = (int)threadIdx.x;
  // (row r, column j), LDS slot r * W + j
  constexpr int ninner = kTriRows * kTriTile;
  int r, j;
  if (tid < ninner) {
    r = tid / kTriTile;
    j = 2 + tid - r * kTriTile;
  } else {
    const int k = tid - ninner;
    r = k >> 2;
    const int m = k & 3;
    j = m < 2 ? m : W - 4 + m;
  }
  const int slot = r * W + j;
  // column before the periodic wrap: -2 .. ctiles * 64 + 1, below 2 * nc
  const int cl = ct * kTriTile - 2 + j;
  const int c = cl < 0 ? cl + nc : (cl >= nc ? cl - nc : cl);
  constexpr bool PER_LANE = ROLE == TRI_MIXED || ROLE == TRI_H;
  const bool active = !PER_LANE || r < kTriRows;
  // rows b0 - 2 .. b0 + 12; a thread past them takes the last row's cell
  const int b = (b0 - 2 + (active ? r : kTriRows - 1) + 2 * nb) % nb;
  const bool in2 = ROLE == TRI_S2 || ROLE == TRI_S3 ||
                   (PER_LANE && active && r >= 1 && r <= kTriRows - 2 && j >= 1 && j <= W - 2);
  const bool in3 = ROLE == TRI_S3 ||
                   (ROLE == TRI_MIXED && r >= 2 && r <= kTriRows - 3 && b0 + r - 2 < nb &&
                    j >= 2 && j < W - 2 && cl < nc);
  const T* src = static_cast<const T*>(a.src);
  const int plane = nb * nc;
  uint32_t off[3][3];
  {
    const int Bs[3] = {b == 0 ? nb - 1 : b - 1, b, b == nb - 1 ? 0 : b + 1};
    const int Cs[3] = {c == 0 ? nc - 1 : c - 1, c, c == nc - 1 ? 0 : c + 1};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) off[i][k] = (uint32_t)((Bs[i] * nc + Cs[k]) * (int)sizeof(T));
  }
  const __amdgpu_buffer_rsrc_t rdst = state_rsrc(a.dst);
  auto qplane = [&](const void* base, int q, int ia) {
    return static_cast<const char*>(base) + ((size_t)q * P + (size_t)ia * plane) * sizeof(T);
  };
  // plane p (a0 - 2 <= p <= a1 + 2) of a periodic axis of na >= 1 planes
  auto wrap_a = [&](int p) {
    int pc = p % na;
    return pc < 0 ? pc + na : pc;
  };
  auto fetch = [&](int p, C (&v)[19]) {
    const int pc = wrap_a(p);
    asm volatile("" : "+v"(off[1][1]));
#pragma unroll
    for (int q = 0; q < 19; ++q)
      v[q] = load_u<L, T, SCALED>(reinterpret_cast<const T*>(qplane(src, q, pc) + off[1][1]), q);
  };
  auto collide = [&](C (&v)[19]) {
    C drho, rho, ua, ub, uc;
    moments<L, C, false, LS>(v, C(0), C(0), C(0), drho, rho, ua, ub, uc);
    collide_for<L, T, C, false, SCALED>(v, a.inv_tau, C(0), C(0), C(0), drho, rho, ua, ub, uc);
  };
  // a step's gather of the 16 directions that cross to another thread: the
  // cell at (r - e_b, j - e_c), which lies in the tile for every cell that
  // runs the step, as slot base sb plus a constant offset (sb = slot - W - 1
  // >= 0 there, so all 16 reads are one register plus immediates)
  auto gather = [&](float (&x)[16][1024], C (&v)[19]) {
    const float* xb = &x[0][0] + (slot - W - 1);
#pragma unroll
    for (int q = 3; q < 19; ++q)
      v[q] = xb[(q - 3) * 1024 + (1 - L::eb[q]) * W + (1 - L::ec[q])];
  };
  // own cell's results of the last two planes: Z (e_a = 0) of the last,
  // Q1 / Q2 (e_a = +1) of the last and the one before
  auto carry = [&](const C (&v)[19], C (&Z)[19], C (&Q1)[19], C (&Q2)[19]) {
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      if (L::ea[q] > 0) {
        Q2[q] = Q1[q];
        Q1[q] = v[q];
      }
      if (L::ea[q] == 0) Z[q] = v[q];
    }
  };
  auto publish = [&](float (&x)[16][1024], const C (&v)[19], const C (&Z)[19], const C (&Q2)[19]) {
#pragma unroll
    for (int q = 3; q < 19; ++q) x[q - 3][slot] = L::ea[q] < 0 ? v[q] : (L::ea[q] == 0 ? Z[q] : Q2[q]);
  };

  C Z1[19], Q11[19], Q21[19], Z2[19], Q12[19], Q22[19], nx[19];
#pragma unroll
  for (int q = 0; q < 19; ++q) Z1[q] = Q11[q] = Q21[q] = Z2[q] = Q12[q] = Q22[q] = C(0);
  fetch(a0 - 2, nx);
  // one plane: step 1 at p, step 2 at p - 1 and, with step 3 (WITH3), step 3
  // at p - 2
  auto plane_at = [&](int p, auto with3) {
    constexpr bool WITH3 = decltype(with3)::value;
    C d[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) d[q] = nx[q];
    fetch(p <= a1 ? p + 1 : a1 + 1, nx);
    collide(d);
    as_stored<L, T, C>(d);
    if (active) publish(x1, d, Z1, Q21);
    // step 2's own-column inputs: rest (plane p - 1), +a (p - 2), -a (p)
    const C o0 = Z1[0], o1 = Q21[1], o2 = d[2];
    carry(d, Z1, Q11, Q21);
    __syncthreads();
    C t0 = C(0), t1 = C(0), t2 = C(0);
    if (ROLE != TRI_S1 && in2) {
      C e[19];
      e[0] = o0;
      e[1] = o1;
      e[2] = o2;
      gather(x1, e);
      collide(e);
      as_stored<L, T, C>(e);
      publish(x2, e, Z2, Q22);
      // step 3's own-column inputs: rest (p - 2), +a (p - 3), -a (p - 1)
      t0 = Z2[0];
      t1 = Q22[1];
      t2 = e[2];
      carry(e, Z2, Q12, Q22);
    }
    __syncthreads();
    if constexpr (WITH3) {
      if (ROLE != TRI_S1 && ROLE != TRI_S2 && ROLE != TRI_H && in3) {
        C g[19];
        g[0] = t0;
        g[1] = t1;
        g[2] = t2;
        gather(x2, g);
        collide(g);
        const uint32_t po = (uint32_t)((p - 2) * plane * (int)sizeof(T));
#pragma unroll
        for (int q = 0; q < 19; ++q)
          buf_store<T>(enc_u<L, T, SCALED>(q, g[q]), rdst, off[1][1],
                       (uint32_t)q * (uint32_t)P * (uint32_t)sizeof(T) + po);
      }
      // keeps the scheduler from mixing the next plane's step 1 into step
      // 3, which needs more registers than a 1024-thread workgroup has
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // planes a0 - 2 .. a0 + 1: step 3 would be at a plane before the chunk
#pragma unroll 1
  for (int p = a0 - 2; p < a0 + 2; ++p) plane_at(p, std::false_type{});
  // The loop's first step 1 consumes the gather in flight anyway; waiting
  // for it here leaves the loop header's memory waits to the loop's own
  // order (gather, then stores). s_waitcnt vmcnt(0), expcnt and lgkmcnt at
  // their maxima.
  __builtin_amdgcn_s_waitcnt(0x0F70);
  for (int p = a0 + 2; p <= a1 + 1; ++p) plane_at(p, std::true_type{});
}

template <class T>
__global__ void __launch_bounds__(1024) k_tri(PairArgs a) {
  __shared__ float x1[16][1024];
  __shared__ float x2[16][1024];
  const int lb = xcd_remap(blockIdx.x, a.nblocks, a.remap);
  const int rest = lb / a.ctiles;
  const int ct = lb - rest * a.ctiles, bt = rest % a.btiles, chunk = rest / a.btiles;
  // fixed roles for the row waves of a c-tile the lattice does not cut
  const int wave = (int)threadIdx.x / 64;
  if (wave < kTriRows && (ct + 1) * kTriTile <= a.nc) {
    const int s3 = wave >= 2 && wave <= kTriRows - 3 && bt * kTriOut + wave - 2 < a.nb;
    const int s2 = wave >= 1 && wave <= kTriRows - 2;
    const int role = __builtin_amdgcn_readfirstlane(s3 ? TRI_S3 : (s2 ? TRI_S2 : TRI_S1));
    if (role == TRI_S3)
      tri_march<T, TRI_S3>(a, ct, bt, chunk, x1, x2);
    else if (role == TRI_S2)
      tri_march<T, TRI_S2>(a, ct, bt, chunk, x1, x2);
    else
      tri_march<T, TRI_S1>(a, ct, bt, chunk, x1, x2);
    return;
  }
  if (wave == kTriRows && (ct + 1) * kTriTile <= a.nc) {
    tri_march<T, TRI_H>(a, ct, bt, chunk, x1, x2);
    return;
  }
  tri_march<T, TRI_MIXED>(a, ct, bt, chunk, x1, x2);
}

}  // namespace mlbm
