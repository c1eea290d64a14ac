// Device-side building blocks shared by every LBM kernel: lattice tables,
// storage traits, the Guo-forced BGK collision, and the row/neighbour index
// arithmetic of the SoA layout.
//
// Internal axes: a state is addressed as [q][ia][ib][ic] with ic contiguous.
// D3Q19 maps the reference's grid axes (0, 1, 2) to (a, b, c); D2Q9 maps its
// axes (0, 1) to (b, c) and has na = 1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlbm {

// ---------------------------------------------------------------------------
// Lattices. Direction order, opposite map and weights follow the reference's
// d3q19()/d2q9() (src/maddening/nodes/lbm.py:90-173).
// ---------------------------------------------------------------------------
template <int Q> struct Lattice;

template <> struct Lattice<19> {
  static constexpr int Q = 19;
  static constexpr int D = 3;
  static constexpr int ea[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int eb[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int ec[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  // weight numerators over 36: 1/3, 1/18 (x6), 1/36 (x12)
  static constexpr int w36[19] = {12, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
};

template <> struct Lattice<9> {
  static constexpr int Q = 9;
  static constexpr int D = 2;
  static constexpr int ea[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int eb[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
  static constexpr int ec[9] = {0, 0, 0, 1, -1, 1, 1, -1, -1};
  static constexpr int opp[9] = {0, 2, 1, 4, 3, 8, 7, 6, 5};
  // 4/9, 1/9 (x4), 1/36 (x4)
  static constexpr int w36[9] = {16, 4, 4, 4, 4, 1, 1, 1, 1};
};

template <class L, class C>
__device__ __forceinline__ constexpr C weight(int q) {
  return C(L::w36[q]) / C(36);
}

// ---------------------------------------------------------------------------
// Storage traits. Every kernel works on deviations d_q = f_q - w_q in its
// compute type C; the trait converts from/to what is stored.
// ---------------------------------------------------------------------------
template <class T> struct Storage;

// Each trait: load (stored value -> deviation), enc (deviation -> stored
// value), put (non-temporal store of a stored value) and store = put(enc).
template <> struct Storage<float> {
  using C = float;
  template <class L>
  static __device__ __forceinline__ float load(const float* p, int q) {
    return *p - weight<L, float>(q);
  }
  template <class L>
  static __device__ __forceinline__ float enc(int q, float d) { return weight<L, float>(q) + d; }
  static __device__ __forceinline__ void put(float* p, float v) { __builtin_nontemporal_store(v, p); }
  template <class L>
  static __device__ __forceinline__ void store(float* p, int q, float d) { put(p, enc<L>(q, d)); }
};

template <> struct Storage<double> {
  using C = double;
  template <class L>
  static __device__ __forceinline__ double load(const double* p, int q) {
    return *p - weight<L, double>(q);
  }
  template <class L>
  static __device__ __forceinline__ double enc(int q, double d) { return weight<L, double>(q) + d; }
  static __device__ __forceinline__ void put(double* p, double v) { __builtin_nontemporal_store(v, p); }
  template <class L>
  static __device__ __forceinline__ void store(double* p, int q, double d) { put(p, enc<L>(q, d)); }
};

// float32 deviation f - w: same 4 bytes as f32, but the rounding of the
// stored value scales with |f - w| instead of with w.
struct f32dev_t {
  float v;
};

template <> struct Storage<f32dev_t> {
  using C = float;
  template <class L>
  static __device__ __forceinline__ float load(const f32dev_t* p, int) {
    return p->v;
  }
  template <class L>
  static __device__ __forceinline__ f32dev_t enc(int, float d) { return f32dev_t{d}; }
  static __device__ __forceinline__ void put(f32dev_t* p, f32dev_t v) {
    __builtin_nontemporal_store(v.v, &p->v);
  }
  template <class L>
  static __device__ __forceinline__ void store(f32dev_t* p, int q, float d) { put(p, enc<L>(q, d)); }
};

// IEEE half holds (f - w) * 2^10: deviations stay below ~1, so the scaled
// values stay far from the 65504 limit while everything above 6e-8 (instead
// of 6e-5) keeps a full 11-bit mantissa. Scaling by a power of two is exact.
constexpr float kHalfScale = 1024.0f;

template <> struct Storage<_Float16> {
  using C = float;
  template <class L>
  static __device__ __forceinline__ float load(const _Float16* p, int) {
    return (float)(*p) * (1.0f / kHalfScale);
  }
  template <class L>
  static __device__ __forceinline__ _Float16 enc(int, float d) { return (_Float16)(d * kHalfScale); }
  static __device__ __forceinline__ void put(_Float16* p, _Float16 v) {
    __builtin_nontemporal_store(v, p);
  }
  template <class L>
  static __device__ __forceinline__ void store(_Float16* p, int q, float d) { put(p, enc<L>(q, d)); }
};

template <> struct Storage<__bf16> {
  using C = float;
  template <class L>
  static __device__ __forceinline__ float load(const __bf16* p, int) {
    return (float)(*p);
  }
  template <class L>
  static __device__ __forceinline__ __bf16 enc(int, float d) { return (__bf16)d; }
  static __device__ __forceinline__ void put(__bf16* p, __bf16 v) { __builtin_nontemporal_store(v, p); }
  template <class L>
  static __device__ __forceinline__ void store(__bf16* p, int q, float d) { put(p, enc<L>(q, d)); }
};

// e_q . (a, b, c) for a lattice vector (components -1, 0 or +1), optionally
// added to `init`. Terms with a zero component are left out and +-1
// products are exact, so the result equals init + e_a a + e_b b + e_c c
// evaluated left to right for finite operands, without the multiplies by
// zero and additions of zero that IEEE arithmetic would otherwise keep.
template <class C>
__device__ __forceinline__ void eterm(int e, C v, C& s, bool& any) {
  if (e == 0) return;
  const C t = e > 0 ? v : -v;
  s = any ? s + t : t;
  any = true;
}
template <class L, class C>
__device__ __forceinline__ C edot(int q, C a, C b, C c) {
  C s = C(0);
  bool any = false;
  eterm(L::ea[q], a, s, any);
  eterm(L::eb[q], b, s, any);
  eterm(L::ec[q], c, s, any);
  return s;
}
template <class L, class C>
__device__ __forceinline__ C edot_from(int q, C init, C a, C b, C c) {
  C s = init;
  bool any = true;
  eterm(L::ea[q], a, s, any);
  eterm(L::eb[q], b, s, any);
  eterm(L::ec[q], c, s, any);
  return s;
}

// 1 / x: v_rcp_f32 (within 1 ulp, x normal) for float, correctly rounded for
// double. Every kernel takes these reciprocals here, so all round alike.
template <class C>
__device__ __forceinline__ C recip(C x) {
  return C(1) / x;
}
template <>
__device__ __forceinline__ float recip<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}

// ---------------------------------------------------------------------------
// Macroscopic moments with the Guo half-force correction
// (lbm.py:209-239): rho = sum f, u = (sum f e + F/2) / max(rho, 1e-10).
// Computed from deviations: rho = 1 + sum d (sum w = 1), sum f e = sum d e.
// ---------------------------------------------------------------------------
template <class L, class C>
__device__ __forceinline__ void moments(const C (&d)[L::Q], C Fa, C Fb, C Fc,
                                        C& drho, C& rho, C& ua, C& ub, C& uc) {
  C s = 0, ja = 0, jb = 0, jc = 0;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    s += d[q];
    if (L::ea[q] > 0) ja += d[q];
    if (L::ea[q] < 0) ja -= d[q];
    if (L::eb[q] > 0) jb += d[q];
    if (L::eb[q] < 0) jb -= d[q];
    if (L::ec[q] > 0) jc += d[q];
    if (L::ec[q] < 0) jc -= d[q];
  }
  drho = s;
  rho = C(1) + s;
  const C inv = recip(rho > C(1e-10) ? rho : C(1e-10));
  ua = (ja + C(0.5) * Fa) * inv;
  ub = (jb + C(0.5) * Fb) * inv;
  uc = (jc + C(0.5) * Fc) * inv;
}

// ---------------------------------------------------------------------------
// BGK collision with the Guo source, in place on deviations
// (lbm.py:934-939 with _equilibrium lbm.py:180-206, _guo_forcing
// lbm.py:307-353):
//   f* = f - (f - feq)/tau + (1 - 1/(2 tau)) w [3 (e - u).F + 9 (e.u)(e.F)]
//   feq - w = w [ (rho - 1) + rho (3 e.u + 4.5 (e.u)^2 - 1.5 u.u) ]
// cs2 = 1/3 gives the exact constants 3, 4.5, 1.5 and 9.
//
// Evaluated once per pair (q, opp q): e.u and e.F change sign between the
// two, so feq splits into a shared even part A = (rho - 1) + rho (4.5 (e.u)^2
// - 1.5 u.u) and an odd part 3 rho e.u, and the source into the even
// 9 (e.u)(e.F) - 3 u.F and the odd 3 e.F. Multiply-adds are written as
// explicit fma so every kernel variant rounds identically.
//
// FIXUP (float32 storage): in exact arithmetic the collision keeps
// sum d = rho - 1 and moves sum d e from j to j + F. Float32 rounding misses
// both by about an ulp per step, which accumulates in rho and u over a run,
// so the miss is projected back onto the conserved moments: adding
// w_q (a + 3 e_q . b) changes sum d by a and sum d e by b, because sum w = 1,
// sum w e = 0 and sum w e e = I/3. (With 16-bit storage the rounding of the
// stored values dominates and the fix-up buys nothing.)
// ---------------------------------------------------------------------------
template <class C>
__device__ __forceinline__ C fmadd(C a, C b, C c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmadd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

template <class L, class C, bool FIXUP>
__device__ __forceinline__ void collide_impl(C (&d)[L::Q], C inv_tau, C Fa, C Fb, C Fc, C drho,
                                             C rho, C ua, C ub, C uc) {
  C ja = 0, jb = 0, jc = 0;
  if constexpr (FIXUP) {
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      if (L::ea[q] > 0) ja += d[q];
      if (L::ea[q] < 0) ja -= d[q];
      if (L::eb[q] > 0) jb += d[q];
      if (L::eb[q] < 0) jb -= d[q];
      if (L::ec[q] > 0) jc += d[q];
      if (L::ec[q] < 0) jc -= d[q];
    }
  }
  const C usq15 = C(1.5) * fmadd(ua, ua, fmadd(ub, ub, uc * uc));
  const C uF3 = C(3) * fmadd(ua, Fa, fmadd(ub, Fb, uc * Fc));
  const C keep = C(1) - inv_tau;
  const C src = C(1) - C(0.5) * inv_tau;
  const C rho3 = C(3) * rho;
  const C rho45 = C(4.5) * rho;
  // rest population: e = 0
  {
    const C w = weight<L, C>(0);
    const C A = fmadd(-rho, usq15, drho);
    d[0] = fmadd(keep, d[0], fmadd(inv_tau * w, A, -(src * w) * uF3));
  }
#pragma unroll
  for (int q = 1; q < L::Q; ++q) {
    if (L::opp[q] < q) continue;  // each pair once, from its first member
    const int m = L::opp[q];
    const C w = weight<L, C>(q);
    const C eu = edot<L, C>(q, ua, ub, uc);
    const C eF = edot<L, C>(q, Fa, Fb, Fc);
    const C A = fmadd(rho45 * eu, eu, fmadd(-rho, usq15, drho));  // even part of (feq - w)/w
    const C B = rho3 * eu;                                         // odd part
    const C sw = src * w;
    const C wt = inv_tau * w;
    const C even = fmadd(wt, A, sw * fmadd(C(9) * eu, eF, -uF3));
    const C odd = fmadd(wt, B, (C(3) * sw) * eF);
    const C dq = d[q];
    const C dm = d[m];
    d[q] = fmadd(keep, dq, even + odd);
    d[m] = fmadd(keep, dm, even - odd);
  }
  if constexpr (FIXUP) {
    C m0 = 0, ma = 0, mb = 0, mc = 0;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      m0 += d[q];
      if (L::ea[q] > 0) ma += d[q];
      if (L::ea[q] < 0) ma -= d[q];
      if (L::eb[q] > 0) mb += d[q];
      if (L::eb[q] < 0) mb -= d[q];
      if (L::ec[q] > 0) mc += d[q];
      if (L::ec[q] < 0) mc -= d[q];
    }
    // targets: rho - 1 (unchanged) and j + F
    const C a0 = drho - m0;
    const C ba = C(3) * ((ja + Fa) - ma);
    const C bb = C(3) * ((jb + Fb) - mb);
    const C bc = C(3) * ((jc + Fc) - mc);
#pragma unroll
    for (int q = 0; q < L::Q; ++q)
      d[q] = fmadd(weight<L, C>(q), edot_from<L, C>(q, a0, ba, bb, bc), d[q]);
  }
}

// Collision for storage T: the float32 fix-up where the stored values are
// float32 (f or f - w), none for float64 (exact enough) or 16-bit storage.
template <class L, class T, class C>
__device__ __forceinline__ void collide_for(C (&d)[L::Q], C inv_tau, C Fa, C Fb, C Fc, C drho,
                                            C rho, C ua, C ub, C uc) {
  constexpr bool FIXUP = sizeof(C) == 4 && sizeof(T) == 4;
  collide_impl<L, C, FIXUP>(d, inv_tau, Fa, Fb, Fc, drho, rho, ua, ub, uc);
}

// ---------------------------------------------------------------------------
// Periodic neighbour rows. For a cell (ia, ib, ic) the index of the cell at
// offset (sa, sb, sc) in {-1,0,1}^3 is row[sa+1][sb+1] + col[sc+1].
// ---------------------------------------------------------------------------
struct Nbr {
  int row[3][3];
  int col[3];
  int cell;
  __device__ __forceinline__ void init(int ia, int ib, int ic, int na, int nb,
                                       int nc) {
    const int A[3] = {ia == 0 ? na - 1 : ia - 1, ia, ia == na - 1 ? 0 : ia + 1};
    const int B[3] = {ib == 0 ? nb - 1 : ib - 1, ib, ib == nb - 1 ? 0 : ib + 1};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) row[i][j] = (A[i] * nb + B[j]) * nc;
    col[0] = ic == 0 ? nc - 1 : ic - 1;
    col[1] = ic;
    col[2] = ic == nc - 1 ? 0 : ic + 1;
    cell = row[1][1] + ic;
  }
  // index of cell x + s * e_q
  template <class L>
  __device__ __forceinline__ int at(int q, int s) const {
    return row[1 + s * L::ea[q]][1 + s * L::eb[q]] + col[1 + s * L::ec[q]];
  }
};

// Map a linear workgroup id to a logical row-block id so that runs of G
// consecutive logical blocks land on one XCD. Workgroups are dealt to the 8
// XCDs round-robin, so physical block b runs on XCD b % 8; the 8 XCDs then
// work on 8 adjacent runs at a time and neighbouring rows share an L2.
// G <= 0 is the identity; G = nblocks / 8 gives each XCD one contiguous
// eighth of the lattice. Blocks past the last whole set of 8 runs keep
// their id.
__device__ __forceinline__ int xcd_remap(int b, int nblocks, int G) {
  if (G <= 0) return b;
  const int span = 8 * G;
  if (b >= (nblocks / span) * span) return b;
  const int slot = b >> 3;
  return (slot / G) * span + (b & 7) * G + slot % G;
}

}  // namespace mlbm
