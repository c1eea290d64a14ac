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
// value), quant (deviation -> the deviation a store and reload gives back,
// for the storage types that hold deviations), put (non-temporal store of a
// stored value) and store = put(enc).
template <> struct Storage<float> {
  using C = float;
  // quant(d) equals a store and reload of d only where Sterbenz holds
  static constexpr bool kQuantExact = false;
  template <class L>
  static __device__ __forceinline__ float load(const float* p, int q) {
    return *p - weight<L, float>(q);
  }
  template <class L>
  static __device__ __forceinline__ float enc(int q, float d) { return weight<L, float>(q) + d; }
  // the deviation a store and reload gives back: (w + d) - w, exact in
  // float32 for |d| < w
  template <class L>
  static __device__ __forceinline__ float quant(int q, float d) {
    return (weight<L, float>(q) + d) - weight<L, float>(q);
  }
  static __device__ __forceinline__ void put(float* p, float v) { __builtin_nontemporal_store(v, p); }
  template <class L>
  static __device__ __forceinline__ void store(float* p, int q, float d) { put(p, enc<L>(q, d)); }
};

template <> struct Storage<double> {
  using C = double;
  static constexpr bool kQuantExact = false;  // no conservation step (float64)
  template <class L>
  static __device__ __forceinline__ double load(const double* p, int q) {
    return *p - weight<L, double>(q);
  }
  template <class L>
  static __device__ __forceinline__ double enc(int q, double d) { return weight<L, double>(q) + d; }
  template <class L>
  static __device__ __forceinline__ double quant(int, double d) { return d; }
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
  static constexpr bool kQuantExact = true;
  template <class L>
  static __device__ __forceinline__ float load(const f32dev_t* p, int) {
    return p->v;
  }
  template <class L>
  static __device__ __forceinline__ f32dev_t enc(int, float d) { return f32dev_t{d}; }
  template <class L>
  static __device__ __forceinline__ float quant(int, float d) { return d; }
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
  static constexpr bool kQuantExact = true;
  static constexpr int kLog2Scale = 10;  // stored value = deviation * 2^10
  template <class L>
  static __device__ __forceinline__ float load(const _Float16* p, int) {
    return (float)(*p) * (1.0f / kHalfScale);
  }
  template <class L>
  static __device__ __forceinline__ _Float16 enc(int, float d) { return (_Float16)(d * kHalfScale); }
  // the deviation a store and reload of d gives back
  template <class L>
  static __device__ __forceinline__ float quant(int, float d) {
    return (float)(_Float16)(d * kHalfScale) * (1.0f / kHalfScale);
  }
  // the same three on deviations kept in stored units (d * 2^10). quant_s
  // rounds D to float32 first, as quant does (it rounds d, then scales it
  // exactly): left to itself the compiler fuses the fma or subtraction that
  // made D with the conversion (v_fma_mixlo_f16), rounding once, straight to
  // half, and the result could differ from quant's by an ulp.
  static __device__ __forceinline__ float load_s(const _Float16* p) { return (float)(*p); }
  static __device__ __forceinline__ _Float16 enc_s(float D) { return (_Float16)D; }
  static __device__ __forceinline__ float quant_s(float D) {
    asm volatile("" : "+v"(D));
    return (float)(_Float16)D;
  }
  static __device__ __forceinline__ void put(_Float16* p, _Float16 v) {
    __builtin_nontemporal_store(v, p);
  }
  template <class L>
  static __device__ __forceinline__ void store(_Float16* p, int q, float d) { put(p, enc<L>(q, d)); }
};

template <> struct Storage<__bf16> {
  using C = float;
  static constexpr bool kQuantExact = true;
  template <class L>
  static __device__ __forceinline__ float load(const __bf16* p, int) {
    return (float)(*p);
  }
  template <class L>
  static __device__ __forceinline__ __bf16 enc(int, float d) { return (__bf16)d; }
  template <class L>
  static __device__ __forceinline__ float quant(int, float d) { return (float)(__bf16)d; }
  static __device__ __forceinline__ void put(__bf16* p, __bf16 v) { __builtin_nontemporal_store(v, p); }
  template <class L>
  static __device__ __forceinline__ void store(__bf16* p, int q, float d) { put(p, enc<L>(q, d)); }
};

// ---------------------------------------------------------------------------
// Stored units. Float16 storage holds deviations times 2^10; a kernel may
// keep its deviations in those units (D = d * 2^kLog2Scale) so that loads,
// the rounding to the stored format and stores need no multiply. Every
// product and sum of the collision is then the exact scaling of its value in
// deviation units (a power of two, far from float32's range limits), so the
// stored bits are the same either way. Storage types without a scale use
// their deviation units (kLog2Scale 0).
// ---------------------------------------------------------------------------
template <class T, class = void>
struct ScaleOf {
  static constexpr int v = 0;
};
template <class T>
struct ScaleOf<T, decltype((void)Storage<T>::kLog2Scale)> {
  static constexpr int v = Storage<T>::kLog2Scale;
};

// 2^k for 0 <= k < 31, as a compile-time float
template <class C, int K>
__host__ __device__ constexpr C pow2() {
  return C(1u << K);
}

// load / rounding to the stored format / encode in stored units (SCALED) or
// deviation units
template <class L, class T, bool SCALED>
__device__ __forceinline__ typename Storage<T>::C load_u(const T* p, int q) {
  if constexpr (SCALED && ScaleOf<T>::v != 0) return Storage<T>::load_s(p);
  else return Storage<T>::template load<L>(p, q);
}
template <class L, class T, bool SCALED, class C>
__device__ __forceinline__ C quant_u(int q, C v) {
  if constexpr (SCALED && ScaleOf<T>::v != 0) return Storage<T>::quant_s(v);
  else return Storage<T>::template quant<L>(q, v);
}
template <class L, class T, bool SCALED, class C>
__device__ __forceinline__ T enc_u(int q, C v) {
  if constexpr (SCALED && ScaleOf<T>::v != 0) return Storage<T>::enc_s(v);
  else return Storage<T>::template enc<L>(q, v);
}

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
// Sums over the populations, each accumulated in q order from its first
// term: sum d, and sum d e along each internal axis.
// ---------------------------------------------------------------------------
template <class L, class C>
__device__ __forceinline__ C mass_sum(const C (&d)[L::Q]) {
  C s = d[0];
#pragma unroll
  for (int q = 1; q < L::Q; ++q) s += d[q];
  return s;
}

template <class L>
__host__ __device__ constexpr int ecomp(int axis, int q) {
  return axis == 0 ? L::ea[q] : (axis == 1 ? L::eb[q] : L::ec[q]);
}

// sum over q != skip of e_q,axis d_q (0 when no term is left)
template <class L, int AXIS, int SKIP, class C>
__device__ __forceinline__ C axis_sum(const C (&d)[L::Q]) {
  C s = C(0);
  bool any = false;
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    if (q == SKIP) continue;
    eterm(ecomp<L>(AXIS, q), d[q], s, any);
  }
  return s;
}

// ---------------------------------------------------------------------------
// Macroscopic moments with the Guo half-force correction
// (lbm.py:209-239): rho = sum f, u = (sum f e + F/2) / max(rho, 1e-10).
// Computed from deviations: rho = 1 + sum d (sum w = 1), sum f e = sum d e.
// ---------------------------------------------------------------------------
//
// HASF = false: the caller's force is zero (uniform, no field). Every force
// term is then an exact zero, so leaving them out changes no value (at most
// the sign of a zero result), and a kernel that knows it has no force saves
// their arithmetic and registers.
//
// LOG2S > 0: d holds deviations in stored units (times 2^LOG2S); the sums
// are scaled back exactly, so the moments are the same bits.
template <class L, class C, bool HASF = true, int LOG2S = 0>
__device__ __forceinline__ void moments(const C (&d)[L::Q], C Fa, C Fb, C Fc,
                                        C& drho, C& rho, C& ua, C& ub, C& uc) {
  constexpr C un = C(1) / pow2<C, LOG2S>();
  const C s = LOG2S ? mass_sum<L, C>(d) * un : mass_sum<L, C>(d);
  const C ja = LOG2S ? axis_sum<L, 0, -1, C>(d) * un : axis_sum<L, 0, -1, C>(d);
  const C jb = LOG2S ? axis_sum<L, 1, -1, C>(d) * un : axis_sum<L, 1, -1, C>(d);
  const C jc = LOG2S ? axis_sum<L, 2, -1, C>(d) * un : axis_sum<L, 2, -1, C>(d);
  drho = s;
  rho = C(1) + s;
  const C inv = recip(rho > C(1e-10) ? rho : C(1e-10));
  if constexpr (HASF) {
    ua = (ja + C(0.5) * Fa) * inv;
    ub = (jb + C(0.5) * Fb) * inv;
    uc = (jc + C(0.5) * Fc) * inv;
  } else {
    ua = ja * inv;
    ub = jb * inv;
    uc = jc * inv;
  }
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
// ---------------------------------------------------------------------------
template <class C>
__device__ __forceinline__ C fmadd(C a, C b, C c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmadd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

template <class L, class C, bool HASF = true>
__device__ __forceinline__ void collide_impl(C (&d)[L::Q], C inv_tau, C Fa, C Fb, C Fc, C drho,
                                             C rho, C ua, C ub, C uc) {
  const C usq15 = C(1.5) * fmadd(ua, ua, fmadd(ub, ub, uc * uc));
  const C keep = C(1) - inv_tau;
  const C rho3 = C(3) * rho;
  const C rho45 = C(4.5) * rho;
  // without a force the source terms below are exact zeros: fma(a, b, +-0)
  // rounds a * b once, as the product alone does
  // rest population: e = 0
  {
    const C w = weight<L, C>(0);
    const C A = fmadd(-rho, usq15, drho);
    if constexpr (HASF) {
      const C uF3 = C(3) * fmadd(ua, Fa, fmadd(ub, Fb, uc * Fc));
      const C src = C(1) - C(0.5) * inv_tau;
      d[0] = fmadd(keep, d[0], fmadd(inv_tau * w, A, -(src * w) * uF3));
    } else {
      d[0] = fmadd(keep, d[0], (inv_tau * w) * A);
    }
  }
#pragma unroll
  for (int q = 1; q < L::Q; ++q) {
    if (L::opp[q] < q) continue;  // each pair once, from its first member
    const int m = L::opp[q];
    const C w = weight<L, C>(q);
    const C eu = edot<L, C>(q, ua, ub, uc);
    const C A = fmadd(rho45 * eu, eu, fmadd(-rho, usq15, drho));  // even part of (feq - w)/w
    const C B = rho3 * eu;                                         // odd part
    const C wt = inv_tau * w;
    C even, odd;
    if constexpr (HASF) {
      const C uF3 = C(3) * fmadd(ua, Fa, fmadd(ub, Fb, uc * Fc));
      const C eF = edot<L, C>(q, Fa, Fb, Fc);
      const C sw = (C(1) - C(0.5) * inv_tau) * w;
      even = fmadd(wt, A, sw * fmadd(C(9) * eu, eF, -uF3));
      odd = fmadd(wt, B, (C(3) * sw) * eF);
    } else {
      even = wt * A;
      odd = wt * B;
    }
    const C dq = d[q];
    const C dm = d[m];
    d[q] = fmadd(keep, dq, even + odd);
    d[m] = fmadd(keep, dm, even - odd);
  }
}

// The direction +1 along internal axis AXIS and 0 along the others
// (D3Q19: q = 1, 3, 5; D2Q9: q = 1, 3 along b, c), -1 when the lattice has
// no such axis.
template <class L>
__host__ __device__ constexpr int axis_dir(int axis) {
  for (int q = 1; q < L::Q; ++q) {
    bool ok = true;
    for (int k = 0; k < 3; ++k) ok = ok && ecomp<L>(k, q) == (k == axis ? 1 : 0);
    if (ok) return q;
  }
  return -1;
}

// Collision for storage T with (float compute) an exact conservation step.
//
// In exact arithmetic the collision keeps sum d = rho - 1 and moves sum d e
// from j to j + F. Rounded results miss both: float32 arithmetic by about an
// ulp per step, and 16-bit storage by the rounding of every stored value
// (about sqrt(Q) times the rounding of one). Over a run the misses add up in
// rho and u, so the stored values are made to meet both targets: every
// population but four is taken as stored (St::quant: the rounding of w + d
// for float32 f, none for f - w in float32) and summed as its pair is
// collided, the direction along +a, +b
// and +c is solved from its axis's momentum target, and the rest population
// from the mass target, each rounded as stored. The miss per step is then
// the rounding of one stored value per moment. The float64 kernels (exact
// enough) skip it.
//
// SCALED: d holds deviations in stored units (ScaleOf<T>); the equilibrium,
// source and force terms are scaled by the same power of two where they meet
// d, so every result is the exact scaling of the unscaled one.
template <class L, class T, class C, bool HASF = true, bool SCALED = false>
__device__ __forceinline__ void collide_for(C (&d)[L::Q], C inv_tau, C Fa, C Fb, C Fc, C drho,
                                            C rho, C ua, C ub, C uc) {
  if constexpr (sizeof(C) != 4) {
    collide_impl<L, C, HASF>(d, inv_tau, Fa, Fb, Fc, drho, rho, ua, ub, uc);
  } else {
    constexpr int LS = SCALED ? ScaleOf<T>::v : 0;
    constexpr C S = pow2<C, LS>();
    constexpr int qa = axis_dir<L>(0), qb = axis_dir<L>(1), qc = axis_dir<L>(2);
    static_assert(qb > 0 && qc > 0, "lattice without +b / +c directions");
    static_assert(L::opp[qb] > qb && L::opp[qc] > qc && (qa < 0 || L::opp[qa] > qa),
                  "solved directions lead their pairs");
    // momentum targets: the momentum before the collision (the sums
    // moments() takes) plus the force
    C ta = axis_sum<L, 0, -1, C>(d), tb = axis_sum<L, 1, -1, C>(d), tc = axis_sum<L, 2, -1, C>(d);
    if constexpr (HASF) {
      ta += LS ? Fa * S : Fa;
      tb += LS ? Fb * S : Fb;
      tc += LS ? Fc * S : Fc;
    }
    const C usq15 = C(1.5) * fmadd(ua, ua, fmadd(ub, ub, uc * uc));
    const C keep = C(1) - inv_tau;
    const C rho3 = C(3) * rho;
    const C rho45 = C(4.5) * rho;
    // sums of the populations taken as stored: mass and e-weighted per axis
    C s0 = C(0), sa = C(0), sb = C(0), sc = C(0);
    bool any0 = false, anya = false, anyb = false, anyc = false;
    auto take = [&](int q, C v) {
      eterm(1, v, s0, any0);
      eterm(L::ea[q], v, sa, anya);
      eterm(L::eb[q], v, sb, anyb);
      eterm(L::ec[q], v, sc, anyc);
    };
#pragma unroll
    for (int q = 1; q < L::Q; ++q) {
      if (L::opp[q] < q) continue;  // each pair once, from its first member
      const int m = L::opp[q];
      const C w = weight<L, C>(q);
      const C eu = edot<L, C>(q, ua, ub, uc);
      const C A = fmadd(rho45 * eu, eu, fmadd(-rho, usq15, drho));  // even part of (feq - w)/w
      const C B = rho3 * eu;                                         // odd part
      // wt and sw carry the scale (exact): even and odd come out in d's units
      const C wt = LS ? (inv_tau * w) * S : inv_tau * w;
      C even, odd;
      if constexpr (HASF) {
        const C uF3 = C(3) * fmadd(ua, Fa, fmadd(ub, Fb, uc * Fc));
        const C eF = edot<L, C>(q, Fa, Fb, Fc);
        const C sw = LS ? ((C(1) - C(0.5) * inv_tau) * w) * S : (C(1) - C(0.5) * inv_tau) * w;
        even = fmadd(wt, A, sw * fmadd(C(9) * eu, eF, -uF3));
        odd = fmadd(wt, B, (C(3) * sw) * eF);
      } else {
        even = wt * A;
        odd = wt * B;
      }
      const bool solved = q == qa || q == qb || q == qc;
      if (!solved) {
        d[q] = quant_u<L, T, SCALED>(q, fmadd(keep, d[q], even + odd));
        take(q, d[q]);
      }
      d[m] = quant_u<L, T, SCALED>(m, fmadd(keep, d[m], even - odd));
      take(m, d[m]);
    }
    if constexpr (qa > 0) {
      d[qa] = quant_u<L, T, SCALED>(qa, ta - sa);
      s0 += d[qa];
    }
    d[qb] = quant_u<L, T, SCALED>(qb, tb - sb);
    d[qc] = quant_u<L, T, SCALED>(qc, tc - sc);
    d[0] = quant_u<L, T, SCALED>(0, (LS ? drho * S : drho) - ((s0 + d[qb]) + d[qc]));
  }
}

// The values a store and reload of collide_for's results give back, for
// kernels that keep them on chip instead (the two-step launches). Where the
// conservation step already rounds them as stored (kQuantExact), they are
// those values.
template <class L, class T, class C>
__device__ __forceinline__ void as_stored(C (&d)[L::Q]) {
  using St = Storage<T>;
  if constexpr (!(St::kQuantExact && sizeof(C) == 4)) {
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
      const T s = St::template enc<L>(q, d[q]);
      d[q] = St::template load<L>(&s, q);
    }
  }
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
