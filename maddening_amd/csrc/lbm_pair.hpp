// Two pull steps per launch for periodic D3Q19 and D2Q9 lattices (temporal
// blocking through LDS). Included by lbm_forward.hip and lbm_batch.hip.
//
// One fused pull launch reads the post-collision state g_n at x - e_q and
// writes g_{n+1}(x) = C(f_{n+1}(x)): 152 B of HBM per cell update in float32.
// k_pair computes g_{n+2} from g_n in one pass, so the state crosses HBM once
// per two updates. A 1024-thread workgroup owns R = 1024 / nc whole rows of
// the contiguous axis c (nc <= 341, so R >= 3; threads past R * nc only join
// the barriers), rows b0 - 1 .. b0 + R - 2 of axis b, and marches along axis
// a over a chunk of planes. Longer rows are cut into c-tiles of tc columns
// with one halo column each side, which compute step 1 only; the rows of a
// workgroup are then those tiles (w = tc + 2 cells, R = 1024 / w):
//
//   plane p:  step 1 at (p, row, c) for all R rows: gather g_n from HBM
//             (the next plane's gather is issued before step 2, so its
//             latency overlaps step 2's arithmetic and stores), collide;
//             step 2 at (p - 1, row, c) for the R - 2 inner rows: gather
//             g_{n+1} from step 1's results, collide, store g_{n+2}.
//
// Step 2 at plane p - 1 needs g_{n+1} from plane p - 2 for the five
// directions with e_a = +1, from plane p - 1 for the nine with e_a = 0 and
// from plane p for the five with e_a = -1. Each thread keeps its own
// cell's values of the last two planes in registers. The 16 directions
// that cross to another thread (all but q = 0, 1, 2, whose source is the
// thread's own column) pass through LDS, written per plane: e_a = -1
// values of plane p, e_a = 0 values of plane p - 1 and e_a = +1 values of
// plane p - 2. Planes alternate between two 64 KB halves, so one barrier per
// plane suffices (256^3 f32dev: 550 us per launch against 597 with one half
// and a second barrier; f32 585 against 656). With whole rows the periodic wrap along c stays
// inside the workgroup; with c-tiles, and along a and b, it is index
// arithmetic.
//
// Each update is the same inlined moments + collision as k_cell's, and step
// 1's results pass through the storage type's encode/decode exactly as a
// store and reload would, so the output equals two k_cell launches bit for
// bit. Halo rows (the first and last of a workgroup) and the plane before
// each chunk are computed twice, by two workgroups; that redundant
// arithmetic and the extra reads it makes are served by the caches.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_common.hpp"
#include "lbm_cell.hpp"

namespace mlbm {

struct PairArgs {
  const void* src;
  void* dst;
  int na, nb, nc;
  int pitch;       // cells from one population plane to the next (>= na * nb * nc);
                   // 19 * pitch < 2^31, so offsets fit 32 bits
  int w;           // cells per workgroup row: nc (whole rows, nc <= 341) or a
                   // c-tile of tc output columns with one halo column each side
  int tc;          // output columns per row: nc, or the tile width (nc >= tc)
  int halo;        // 0 for whole rows, 1 for c-tiles
  int ctiles;      // 1, or ceil(nc / tc)
  int tb;          // output rows per workgroup = 1024 / w - 2
  int btiles;      // ceil(nb / tb)
  int la;          // planes per chunk
  int nblocks;     // ctiles * btiles * ceil(na / la)
  int remap;       // xcd_remap run length (0: identity)
  float inv_tau, Fa, Fb, Fc;
  // Read by the general launches (GEN) only: the reference wall rule at
  // mask != 0, per-cell force (*G, 3) and tau (*G) fields, Zou-He faces;
  // each pointer nullable. rho/u/p: the moments a PAIR_MOMENTS launch writes.
  const uint8_t* mask;
  const float* force_field;
  const float* tau_field;
  float* rho;
  float* u;
  float* p;
  int zh_n;
  int zh_axis[2];  // internal axis 0 (a), 1 (b), 2 (c)
  int zh_side[2];  // 0: first layer, 1: last layer
  float zh_rho[2];
};

// What the two halves of a launch do. Step 1 reads either the post-collision
// state g_{n-1} at x - e_q (PAIR_GATHER: streaming, then the wall rule and
// faces) or the pre-collision state f_n at x (PAIR_LOCAL: an aligned read,
// the first launch of a sequence), and collides. Step 2 gathers from step
// 1's results, applies the wall rule and faces, and either collides and
// stores g (PAIR_COLLIDE) or stores the streamed state f with its moments
// (PAIR_MOMENTS: the last launch of a sequence, or, after PAIR_LOCAL, one
// whole LBMNode.update in one launch).
enum { PAIR_GATHER = 0, PAIR_LOCAL = 1 };
enum { PAIR_COLLIDE = 0, PAIR_MOMENTS = 1 };

// Reference wall rule (lbm.py:944-948) after a gather: a wall cell's slot q
// takes the value streamed into slot opp(q), i.e. the two slots of every
// opposite pair are exchanged.
template <class L, class C>
__device__ __forceinline__ void swap_pairs(C (&d)[L::Q]) {
#pragma unroll
  for (int q = 1; q < L::Q; ++q) {
    if (L::opp[q] < q) continue;
    const C t = d[q];
    d[q] = d[L::opp[q]];
    d[L::opp[q]] = t;
  }
}

// Per-cell inputs of a general launch fetched with the cell's populations
// and carried to step 2. A force field is read where it is used instead
// (three more values per carried plane would spill registers).
struct PairAux {
  uint32_t m;  // mask byte (0: fluid)
  float tau;
};

// Forms of the march, from the leanest: periodic without a force (no force
// terms in the arithmetic), periodic with a uniform force, the general form
// (reference walls, tau and force fields), and the general form with Zou-He
// faces. Each form is its own instantiation, so a form's registers are not
// raised by what only a later form needs.
enum { PAIR_PLAIN = 0, PAIR_FORCE = 1, PAIR_GEN = 2, PAIR_GEN_FACES = 3 };

// Buffer stores for the plain and general marches: one resource for the
// output array (its base, a 4 GiB range, untyped dwords) and two 32-bit
// offsets, the cell's place in a plane (per thread, fixed through the march)
// and the population and plane (uniform), where a 64-bit address costs a
// vector add per store. 256^3 per two-step launch: float16 380 -> 366 us,
// bfloat16 351 -> 331, the 32-bit types unchanged (496-499 us); with fixed
// wave roles, the config-4 sphere channel (general form) 0.0841-0.0846 ->
// 0.0818-0.0824 ms per step. The uniform-force form spills with them (573 ->
// 605 us at 256^3) and keeps 64-bit addresses. The same for the gather was
// slower (f32dev 514 us, float16 478), so it keeps scalar plane bases with
// 32-bit per-thread offsets. The launch uses it when the array's 19
// populations and a plane end within 4 GiB of its base.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t state_rsrc(const void* p) {
  // word 3 of a gfx9 buffer resource for untyped dword access
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xFFFFFFFF,
                                           0x00020000);
}
// non-temporal store (cache policy bit 1: slc / nt), as St::put
template <class T>
__device__ __forceinline__ void buf_store(T x, __amdgpu_buffer_rsrc_t r, uint32_t v, uint32_t s) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 2, "4- or 2-byte storage");
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, (int)v, (int)s, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, x), r, (int)v, (int)s, 2);
}

// What a wave does in step 2. PAIR_MIXED: decided per lane. When whole rows
// fill whole waves (w a multiple of 64), every lane of a wave is in the same
// row, so the wave is either an inner row (PAIR_INNER: step 2 and its stores
// on every plane) or a halo row or a thread past the rows (PAIR_HALO: step 1
// only). A fixed role keeps the stores out of a branch: the memory counter
// then counts a known number of stores, and the wait for the next plane's
// gather at the end of an iteration no longer waits for the stores the
// iteration just issued.
enum { PAIR_MIXED = 0, PAIR_INNER = 1, PAIR_HALO = 2 };

// The march of one workgroup over c-tile ct, b-tile bt and a-chunk `chunk`
// of lattice a; xx is the workgroup's exchange buffer, two halves of 64 KB. With whole rows
// (w == nc) the periodic wrap along c stays inside the workgroup; with
// c-tiles the halo columns compute step 1 only, and step 2 of the tc inner
// columns reads them. BUF: stores through a buffer resource (the plain and
// general forms).
template <class T, int S1, int S2, int FORM, int ROLE = PAIR_MIXED, bool BUF = false>
__device__ __forceinline__ void pair_march(const PairArgs& a, int ct, int bt, int chunk,
                                           float (&xx)[2][16][1024]) {
  using L = Lattice<19>;
  using St = Storage<T>;
  using C = float;
  static_assert(sizeof(typename Storage<T>::C) == 4, "float compute only");
  constexpr bool GEN = FORM >= PAIR_GEN;
  constexpr bool FACES = FORM == PAIR_GEN_FACES;
  constexpr bool HASF = FORM != PAIR_PLAIN;
  // LEAN (the plain form, whose registers allow it): scalar plane bases
  // with per-thread 32-bit offsets for the gather, issued before step 1's
  // arithmetic. The other forms keep 64-bit per-lane addresses and gather
  // after the barrier, which fits them in 128 registers.
  constexpr bool LEAN = FORM == PAIR_PLAIN;
  // float16 storage: deviations in stored units (exact; no multiply per
  // load, rounding or store), except in the faces form, whose face map
  // works on deviations
  constexpr bool SCALED = ScaleOf<T>::v != 0 && !FACES;
  constexpr int LS = SCALED ? ScaleOf<T>::v : 0;
  const int a0 = chunk * a.la;
  const int a1 = a0 + a.la < a.na ? a0 + a.la : a.na;  // output planes [a0, a1)
  const int b0 = bt * a.tb;
  const int nc = a.nc, na = a.na, nb = a.nb, P = a.pitch, w = a.w;
  const int tid = (int)threadIdx.x;
  // (row r, column j) of the workgroup's rows; its LDS slot is r * w + j.
  // Whole rows: thread tid is slot tid. C-tiles: the first R * tc threads are
  // the inner columns, row by row, so each wave stores whole aligned lines;
  // the next 2 R threads are the halo columns (j = 0 and w - 1 of each row).
  int r, j;
  if (a.halo) {
    const int ninner = (a.tb + 2) * a.tc;
    if (tid < ninner) {
      r = tid / a.tc;
      j = 1 + tid - r * a.tc;
    } else {
      const int k = tid - ninner;
      r = k >> 1;
      j = (k & 1) ? w - 1 : 0;
    }
  } else {
    r = tid / w;
    j = tid - r * w;
  }
  const int slot = r * w + j;
  // column before the periodic wrap: -1 .. ctiles * tc, below 2 * nc as nc >= tc
  const int cl = ct * a.tc - a.halo + j;
  const int c = cl < 0 ? cl + nc : (cl >= nc ? cl - nc : cl);
  const bool active = ROLE == PAIR_INNER || r < a.tb + 2;
  // a thread past the workgroup's rows takes the last row's cell: its
  // (unused) gathers then hit lines the workgroup reads anyway
  const int b = (b0 - 1 + (active ? r : a.tb + 1) + nb) % nb;
  const bool inner = ROLE == PAIR_INNER ||
                     (ROLE == PAIR_MIXED && r >= 1 && r <= a.tb && b0 + r - 1 < nb &&
                      j >= a.halo && j < a.halo + a.tc && cl < nc);
  const T* src = static_cast<const T*>(a.src);
  T* dst = static_cast<T*>(a.dst);
  const int plane = nb * nc;
  // LEAN addresses: plane offsets along a are the same for the whole
  // workgroup, so each population's base pointer for a plane is uniform
  // (scalar registers), and a thread adds one of nine 32-bit byte offsets
  // that stay fixed through the march: off[sb + 1][sc + 1] for the cell
  // (b + sb, c + sc).
  uint32_t off[3][3];
  if constexpr (LEAN) {
    const int Bs[3] = {b == 0 ? nb - 1 : b - 1, b, b == nb - 1 ? 0 : b + 1};
    const int Cs[3] = {c == 0 ? nc - 1 : c - 1, c, c == nc - 1 ? 0 : c + 1};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) off[i][k] = (uint32_t)((Bs[i] * nc + Cs[k]) * (int)sizeof(T));
  }
  const __amdgpu_buffer_rsrc_t rdst = state_rsrc(a.dst);
  // population q of plane ia (0 <= ia < na), as a uniform byte pointer
  auto qplane = [&](const void* base, int q, int ia) {
    return static_cast<const char*>(base) + ((size_t)q * P + (size_t)ia * plane) * sizeof(T);
  };

  // step 1's input at plane p (-1 <= p <= na): the populations, and for a
  // general launch the cell's mask byte and tau (loaded with them, used
  // after they land)
  auto fetch = [&](int p, C (&v)[19], PairAux& au) {
    const int pc = p < 0 ? p + na : (p >= na ? p - na : p);
    if constexpr (LEAN) {
      // keep the offsets 32-bit here: hoisted out of the loop, base + offset
      // would become one 64-bit vector pointer per source, held in registers
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) asm volatile("" : "+v"(off[i][k]));
      const int pl[3] = {pc == 0 ? na - 1 : pc - 1, pc, pc == na - 1 ? 0 : pc + 1};
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const char* at =
            S1 == PAIR_LOCAL ? qplane(src, q, pc) + off[1][1]
                             : qplane(src, q, pl[1 - L::ea[q]]) + off[1 - L::eb[q]][1 - L::ec[q]];
        v[q] = load_u<L, T, SCALED>(reinterpret_cast<const T*>(at), q);
      }
    } else {
      Nbr n;
      n.init(pc, b, c, na, nb, nc);
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const int at = S1 == PAIR_LOCAL ? n.cell : n.template at<L>(q, -1);
        v[q] = load_u<L, T, SCALED>(src + q * P + at, q);
      }
      if constexpr (GEN) {
        au.m = a.mask ? a.mask[n.cell] : 0u;
        au.tau = a.tau_field ? a.tau_field[n.cell] : 0.0f;
      }
    }
  };
  auto inv_tau_of = [&](const PairAux& au) {
    return GEN && a.tau_field ? recip(au.tau) : a.inv_tau;
  };
  auto force_at = [&](int cell, C& Fa, C& Fb, C& Fc) {
    Fa = a.Fa;
    Fb = a.Fb;
    Fc = a.Fc;
    if (GEN && a.force_field) {
      Fa = a.force_field[3 * cell + 0];
      Fb = a.force_field[3 * cell + 1];
      Fc = a.force_field[3 * cell + 2];
    }
  };
  // after a gather: the reference wall rule at wall cells, faces elsewhere
  auto walls_faces = [&](C (&v)[19], const PairAux& au, int ia) {
    if constexpr (GEN) {
      if (au.m) swap_pairs<L, C>(v);
      else if constexpr (FACES) apply_faces<L, C>(a, ia, b, c, v);
    }
  };

  // own column's step-1 values: Z of plane p - 1 (e_a = 0), P1 / P2 of planes
  // p - 1 / p - 2 (e_a = +1); only the entries of those classes are used
  C Z[19], P1[19], P2[19], nx[19];
  PairAux nau = {0u, 0.0f}, cur = nau, prev = nau;
#pragma unroll
  for (int q = 0; q < 19; ++q) Z[q] = P1[q] = P2[q] = C(0);
  if (LEAN || active) fetch(a0 - 1, nx, nau);
  // step 1 at plane p on the fetched values. LEAN: the next plane's gather
  // is issued first, so it overlaps this plane's arithmetic; on the last
  // plane it repeats plane a1, and every thread issues it (those past the
  // workgroup's rows load the last row's cells again), so every plane
  // issues the same loads.
  auto step1 = [&](int p, C (&d)[19]) {
#pragma unroll
    for (int q = 0; q < 19; ++q) d[q] = nx[q];
    if constexpr (GEN) cur = nau;
    if constexpr (LEAN) fetch(p < a1 ? p + 1 : a1, nx, nau);
    if constexpr (S1 == PAIR_GATHER) walls_faces(d, cur, (p + na) % na);
    C drho, rho, ua, ub, uc, Fa, Fb, Fc;
    force_at((((p + na) % na) * nb + b) * nc + c, Fa, Fb, Fc);
    moments<L, C, HASF, LS>(d, Fa, Fb, Fc, drho, rho, ua, ub, uc);
    collide_for<L, T, C, HASF, SCALED>(d, inv_tau_of(cur), Fa, Fb, Fc, drho, rho, ua, ub, uc);
    as_stored<L, T, C>(d);  // the values a store and reload would give
  };
  // the own column's values of this plane become those of the last one
  auto carry = [&](const C (&d)[19]) {
    if constexpr (GEN) prev = cur;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      if (L::ea[q] > 0) {
        P2[q] = P1[q];
        P1[q] = d[q];
      }
      if (L::ea[q] == 0) Z[q] = d[q];
    }
  };
  // Planes a0 - 1 and a0 take step 1 only (step 2 on them would write
  // planes outside the chunk). With a fixed role they run before the loop,
  // so the loop body has no branch around step 2.
  constexpr bool PEEL = ROLE != PAIR_MIXED;
  if constexpr (PEEL) {
#pragma unroll 1
    for (int p = a0 - 1; p <= a0; ++p) {
      C d[19];
      step1(p, d);
      if constexpr (!LEAN) {
        if (p < a1 && active) fetch(p + 1, nx, nau);
      }
      carry(d);
    }
    // The loop's first step 1 consumes the gather in flight anyway. Waiting
    // for it here means the loop header's memory waits are set by the loop's
    // own order (gather, then stores) alone; merged with this path, whose
    // newest operations are loads, they would wait for the stores too.
    // s_waitcnt vmcnt(0), expcnt and lgkmcnt left at their maxima.
    __builtin_amdgcn_s_waitcnt(0x0F70);
  }
  for (int p = PEEL ? a0 + 1 : a0 - 1; p <= a1; ++p) {
    C d[19];
    step1(p, d);
    const bool out = PEEL || p > a0;
    // planes alternate between the two halves of the exchange buffer, so the
    // one barrier per plane both publishes this plane's values and retires
    // the reads of the half written next
    float (&x)[16][1024] = xx[p & 1];
    if (out && active) {
#pragma unroll
      for (int q = 3; q < 19; ++q)
        x[q - 3][slot] = L::ea[q] < 0 ? d[q] : (L::ea[q] == 0 ? Z[q] : P2[q]);
    }
    __syncthreads();
    if constexpr (!LEAN) {
      if (p < a1 && active) fetch(p + 1, nx, nau);
    }
    // step 2 at plane p - 1: gather g_{n+1} from step 1's results
    auto gather2 = [&](C (&e)[19]) {
      e[0] = Z[0];   // (p - 1, b, c)
      e[1] = P2[1];  // (p - 2, b, c)
      e[2] = d[2];   // (p, b, c)
#pragma unroll
      for (int q = 3; q < 19; ++q) {
        // the row's neighbour column: periodic inside whole rows; an inner
        // column of a c-tile (j in 1 .. tc) has both neighbours in its tile
        int cc = j - L::ec[q];
        cc = cc < 0 ? cc + w : (cc >= w ? cc - w : cc);
        e[q] = x[q - 3][(r - L::eb[q]) * w + cc];
      }
      walls_faces(e, prev, p - 1);
    };
    if (ROLE != PAIR_HALO && out && inner) {
      C e[19];
      gather2(e);
      const int cell = ((p - 1) * nb + b) * nc + c;
      C drho, rho, ua, ub, uc, Fa, Fb, Fc;
      force_at(cell, Fa, Fb, Fc);
      moments<L, C, HASF, LS>(e, Fa, Fb, Fc, drho, rho, ua, ub, uc);
      if constexpr (S2 == PAIR_MOMENTS) {
        // the stream + moments launch's outputs (lbm.py:970-976)
        if (a.rho) a.rho[cell] = rho;
        if (a.p) a.p[cell] = rho * C(1.0 / 3.0);
        if (a.u) {
          const bool wall = GEN && prev.m;
          a.u[3 * cell + 0] = wall ? C(0) : ua;
          a.u[3 * cell + 1] = wall ? C(0) : ub;
          a.u[3 * cell + 2] = wall ? C(0) : uc;
        }
      } else {
        collide_for<L, T, C, HASF, SCALED>(e, inv_tau_of(prev), Fa, Fb, Fc, drho, rho, ua, ub, uc);
      }
      if constexpr (BUF) {
        const uint32_t po = (uint32_t)((p - 1) * plane * (int)sizeof(T));
        const uint32_t co = LEAN ? off[1][1] : (uint32_t)((b * nc + c) * (int)sizeof(T));
#pragma unroll
        for (int q = 0; q < 19; ++q)
          buf_store<T>(enc_u<L, T, SCALED>(q, e[q]), rdst, co,
                       (uint32_t)q * (uint32_t)P * (uint32_t)sizeof(T) + po);
      } else {
#pragma unroll
        for (int q = 0; q < 19; ++q) St::put(dst + q * P + cell, enc_u<L, T, SCALED>(q, e[q]));
      }
    }
    // keeps the scheduler from mixing the next plane's step 1 into step 2,
    // which needs more than the 128 registers a 1024-thread workgroup has
    if constexpr (PEEL) __builtin_amdgcn_sched_barrier(0);
    carry(d);
  }
}

// The same two steps with every step-1 value held in LDS rings instead of
// registers (152 KB: e_a = -1 values for one plane, e_a = 0 for two, e_a = +1
// for three), for callers that need registers of their own: the batch
// kernel reads each workgroup's lattice from a table, and with the register
// carry above it spilled inside the plane loop, where every reload waits for
// the prefetched gather (5.06 ms per launch on the benchmark batch). Same
// per-cell arithmetic, same bits.
struct PairRing {
  float v[38][1024];  // rows: M 0-4 | Z 5 + 9 * (t & 1) | P 23 + 5 * (t % 3)
};

__host__ __device__ constexpr int ring_class_slot(int q) {
  // M: 2, 8, 10, 12, 14; Z: 0, 3, 4, 5, 6, 15, 16, 17, 18; P: 1, 7, 9, 11, 13
  return q == 2 ? 0 : q == 8 ? 1 : q == 10 ? 2 : q == 12 ? 3 : q == 14 ? 4 :
         q == 0 ? 0 : q == 3 ? 1 : q == 4 ? 2 : q == 5 ? 3 : q == 6 ? 4 : q == 15 ? 5 :
         q == 16 ? 6 : q == 17 ? 7 : q == 18 ? 8 :
         q == 1 ? 0 : q == 7 ? 1 : q == 9 ? 2 : q == 11 ? 3 : 4;
}

// ring row of direction q written by step-1 iteration t
template <class L>
__device__ __forceinline__ int ring_row(int q, int t) {
  return L::ea[q] < 0 ? ring_class_slot(q)
                      : (L::ea[q] == 0 ? 5 + (t & 1) * 9 + ring_class_slot(q)
                                       : 23 + (t % 3) * 5 + ring_class_slot(q));
}

template <class T>
__device__ __forceinline__ void pair_march_ring(const PairArgs& a, int bt, int chunk,
                                                PairRing& x) {
  using L = Lattice<19>;
  using St = Storage<T>;
  using C = float;
  static_assert(sizeof(typename Storage<T>::C) == 4, "float compute only");
  const int a0 = chunk * a.la;
  const int a1 = a0 + a.la < a.na ? a0 + a.la : a.na;  // output planes [a0, a1)
  const int b0 = bt * a.tb;
  const int nc = a.nc, na = a.na, nb = a.nb, P = a.pitch;
  const int tid = (int)threadIdx.x;
  const int r = tid / nc;
  const int c = tid - r * nc;
  const bool active = r < a.tb + 2;
  const int b = (b0 - 1 + r + nb) % nb;
  const bool inner = r >= 1 && r <= a.tb && b0 + r - 1 < nb;
  const T* src = static_cast<const T*>(a.src);
  T* dst = static_cast<T*>(a.dst);
  auto fetch = [&](int p, C (&v)[19]) {
    Nbr n;
    n.init((p + na) % na, b, c, na, nb, nc);
#pragma unroll
    for (int q = 0; q < 19; ++q)
      v[q] = St::template load<L>(src + q * P + n.template at<L>(q, -1), q);
  };
  C nx[19];
  if (active) fetch(a0 - 1, nx);
  for (int p = a0 - 1, t = 0; p <= a1; ++p, ++t) {
    if (active) {
      C d[19];
#pragma unroll
      for (int q = 0; q < 19; ++q) d[q] = nx[q];
      if (p < a1) fetch(p + 1, nx);
      C drho, rho, ua, ub, uc;
      moments<L, C>(d, a.Fa, a.Fb, a.Fc, drho, rho, ua, ub, uc);
      collide_for<L, T, C>(d, a.inv_tau, a.Fa, a.Fb, a.Fc, drho, rho, ua, ub, uc);
      as_stored<L, T, C>(d);  // the values a store and reload would give
#pragma unroll
      for (int q = 0; q < 19; ++q) x.v[ring_row<L>(q, t)][tid] = d[q];
    }
    __syncthreads();
    if (p > a0 && inner) {
      // step 2 at (p - 1, b, c): e_a = -1 from plane p (iteration t), 0 from
      // p - 1 (t - 1: (t - 1) & 1 == (t + 1) & 1), +1 from p - 2 (t - 2:
      // (t - 2) % 3 == (t + 1) % 3)
      C e[19];
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        int cc = c - L::ec[q];  // periodic along c, inside the workgroup's rows
        cc = cc < 0 ? cc + nc : (cc >= nc ? cc - nc : cc);
        e[q] = x.v[ring_row<L>(q, L::ea[q] < 0 ? t : t + 1)][(r - L::eb[q]) * nc + cc];
      }
      C drho, rho, ua, ub, uc;
      moments<L, C>(e, a.Fa, a.Fb, a.Fc, drho, rho, ua, ub, uc);
      collide_for<L, T, C>(e, a.inv_tau, a.Fa, a.Fb, a.Fc, drho, rho, ua, ub, uc);
      const int cell = ((p - 1) * nb + b) * nc + c;
#pragma unroll
      for (int q = 0; q < 19; ++q) St::template store<L>(dst + q * P + cell, q, e[q]);
    }
    __syncthreads();
  }
}

template <class T, int S1, int S2, int FORM, bool BUF = false>
__global__ void __launch_bounds__(1024) k_pair(PairArgs a) {
  __shared__ float x[2][16][1024];
  const int lb = xcd_remap(blockIdx.x, a.nblocks, a.remap);
  const int rest = lb / a.ctiles;
  const int ct = lb - rest * a.ctiles, bt = rest % a.btiles, chunk = rest / a.btiles;
  if constexpr (FORM <= PAIR_GEN && S2 == PAIR_COLLIDE) {
    // whole rows in whole waves: one role per wave (the barrier count per
    // plane is the same in both roles)
    if (!a.halo && a.w % 64 == 0) {
      const int r = (int)threadIdx.x / a.w;
      const bool in = r >= 1 && r <= a.tb && bt * a.tb + r - 1 < a.nb;
      if (__builtin_amdgcn_readfirstlane(in ? 1 : 0))
        pair_march<T, S1, S2, FORM, PAIR_INNER, BUF>(a, ct, bt, chunk, x);
      else
        pair_march<T, S1, S2, FORM, PAIR_HALO, BUF>(a, ct, bt, chunk, x);
      return;
    }
  }
  pair_march<T, S1, S2, FORM, PAIR_MIXED, BUF>(a, ct, bt, chunk, x);
}

// The form a launch needs (PAIR_PLAIN .. PAIR_GEN_FACES).
inline int pair_form(const PairArgs& a) {
  if (a.zh_n) return PAIR_GEN_FACES;
  if (a.mask || a.force_field || a.tau_field) return PAIR_GEN;
  return a.Fa != 0.0f || a.Fb != 0.0f || a.Fc != 0.0f ? PAIR_FORCE : PAIR_PLAIN;
}

// ---------------------------------------------------------------------------
// D2Q9: two pull steps per launch. Rows along c are cut into tiles of 960
// output cells (15 waves, so stores stay 64-cell aligned) with 32-cell halos
// on either side; a 1024-thread workgroup marches along b over a chunk of
// rows. Step 1 runs on all 1024 columns (the halos wrap periodically), step
// 2 on the 960 inner ones, reading its left and right neighbours' step-1
// values from LDS (two halves of 24 KB: the six directions with e_c != 0).
// Same structure and same bits as the D3Q19 march above, with rows in place
// of planes. It takes 68 VGPRs (7 waves per SIMD: one 16-wave workgroup per
// CU); capped at 64 so that two fit, it spilled 3 and was no faster (8192^2:
// 133.4k against 133.5k MLUPS, 2048^2: 81.7k against 85.0k). With 16-bit storage each lane
// moves 2-byte values, and it is 2.2x slower than the row kernels that move
// 2-4 cells per lane (8192^2 f16dev: 71.6k against 158.8k MLUPS), so it runs
// there only under LBM_FLAG_PAIR_FORCE.
// ---------------------------------------------------------------------------
struct Pair9Args {
  const void* src;
  void* dst;
  int nb, nc;
  int pitch;    // cells from one population plane to the next; 9 * pitch < 2^31
  int ctiles;   // ceil(nc / 960)
  int lb;       // rows per chunk
  int nblocks;  // ctiles * ceil(nb / lb)
  float inv_tau, Fb, Fc;
};

template <class T>
__global__ void __launch_bounds__(1024) k_pair9(Pair9Args a) {
  using L = Lattice<9>;
  using St = Storage<T>;
  using C = float;
  static_assert(sizeof(typename Storage<T>::C) == 4, "float compute only");
  __shared__ float xx[2][6][1024];  // q = 3 .. 8, rows alternating between halves
  const int ct = (int)blockIdx.x % a.ctiles;
  const int b0 = ((int)blockIdx.x / a.ctiles) * a.lb;
  const int b1 = b0 + a.lb < a.nb ? b0 + a.lb : a.nb;  // output rows [b0, b1)
  const int nb = a.nb, nc = a.nc, P = a.pitch;
  const int tid = (int)threadIdx.x;
  const int cl = ct * 960 - 32 + tid;   // column before the periodic wrap
  const int c = ((cl % nc) + nc) % nc;
  const bool inner = tid >= 32 && tid < 992 && cl < nc;
  const T* src = static_cast<const T*>(a.src);
  T* dst = static_cast<T*>(a.dst);
  auto fetch = [&](int p, C (&v)[9]) {
    Nbr n;
    n.init(0, (p + nb) % nb, c, 1, nb, nc);
#pragma unroll
    for (int q = 0; q < 9; ++q)
      v[q] = St::template load<L>(src + q * P + n.template at<L>(q, -1), q);
  };
  // own column's step-1 values: Z of row p - 1 (e_b = 0), P1 / P2 of rows
  // p - 1 / p - 2 (e_b = +1); only the entries of those classes are used
  C Z[9], P1[9], P2[9], nx[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) Z[q] = P1[q] = P2[q] = C(0);
  fetch(b0 - 1, nx);
  for (int p = b0 - 1; p <= b1; ++p) {
    C d[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) d[q] = nx[q];
    {
      C drho, rho, ua, ub, uc;
      moments<L, C>(d, C(0), a.Fb, a.Fc, drho, rho, ua, ub, uc);
      collide_for<L, T, C>(d, a.inv_tau, C(0), a.Fb, a.Fc, drho, rho, ua, ub, uc);
      as_stored<L, T, C>(d);  // the values a store and reload would give
    }
    const bool out = p > b0;
    // rows alternate between the two halves, so one barrier per row both
    // publishes this row's values and retires the reads of the half
    // written next (8192^2 f32dev 956 against 969 us per launch with one
    // half and two barriers, f32 937 against 957, 4096^2 254 against 264,
    // 2048^2 83 against 88; issuing the next row's gather before this row's
    // arithmetic as well was faster at 2048^2 and 4096^2 but slower at
    // 8192^2, 1017 against 969)
    float (&x)[6][1024] = xx[p & 1];
    if (out) {
#pragma unroll
      for (int q = 3; q < 9; ++q)
        x[q - 3][tid] = L::eb[q] < 0 ? d[q] : (L::eb[q] == 0 ? Z[q] : P2[q]);
    }
    __syncthreads();
    if (p < b1) fetch(p + 1, nx);
    if (out && inner) {
      C e[9];
      e[0] = Z[0];   // (p - 1, c)
      e[1] = P2[1];  // e_b = +1: (p - 2, c)
      e[2] = d[2];   // e_b = -1: (p, c)
#pragma unroll
      for (int q = 3; q < 9; ++q) e[q] = x[q - 3][tid - L::ec[q]];
      C drho, rho, ua, ub, uc;
      moments<L, C>(e, C(0), a.Fb, a.Fc, drho, rho, ua, ub, uc);
      collide_for<L, T, C>(e, a.inv_tau, C(0), a.Fb, a.Fc, drho, rho, ua, ub, uc);
      const int cell = (p - 1) * nc + c;
#pragma unroll
      for (int q = 0; q < 9; ++q) St::template store<L>(dst + q * P + cell, q, e[q]);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      if (L::eb[q] > 0) {
        P2[q] = P1[q];
        P1[q] = d[q];
      }
      if (L::eb[q] == 0) Z[q] = d[q];
    }
  }
}

}  // namespace mlbm
