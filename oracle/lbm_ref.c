/* C restatement of the reference LBM step: a second ORACLE.
 *
 * Test infrastructure only. Nothing in the product (maddening_amd) links,
 * loads or calls this file or the library built from it. Only tests/ and
 * bench.py's CPU-baseline leg use it, as a checker and as the multi-core CPU
 * timing of the reference algorithm.
 *
 * It restates one LBMNode.update of MADDENING
 * (src/maddening/nodes/lbm.py:899-986) in the reference's layout: the
 * state is (*grid_shape, Q) with the populations of a cell contiguous. The
 * step order is the reference's:
 *   1. moments with the F/2 correction         lbm.py:209-239, 934
 *   2. BGK towards feq, plus the Guo source    lbm.py:180-206, 307-353, 937-939
 *   3. periodic streaming, f_q(x) <- f*_q(x - e_q)   lbm.py:242-267
 *   4. wall rule: node swap f_q <- f_opp(q) at wall cells   lbm.py:944-948
 *      (or, as an extension, link-wise halfway bounce-back with a moving wall)
 *   5. Zou-He pressure faces                    lbm.py:425-524, 950-968
 *   6. output moments, u = 0 at walls, p = rho cs2   lbm.py:970-976
 * The file is compiled twice over through REAL (double and float), giving
 * lbm_oracle_update_f64 / _f32 and lbm_oracle_run_f64 / _f32. Cells are
 * split over OpenMP threads; every cell's arithmetic is serial and written
 * in the order of the reference's expressions.
 *
 * Pinned by tests/test_oracle_c.py against the NumPy restatement
 * (oracle/lbm_numpy.py, itself pinned by closed forms and invariants) and
 * against the committed golden fixtures.
 */
#ifndef REAL

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* lbm.py:92-135 (D3Q19) and lbm.py:144-167 (D2Q9), in the reference's order */
static const int E19[19][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
static const int E9[9][3] = {{0, 0, 0}, {1, 0, 0},  {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                             {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}};

typedef struct {
  int Q, D;
  long long n[3];    /* grid extents; n[2] = 1 for D2Q9 */
  long long ncell;
  int e[19][3];
  int opp[19];
  double w[19];
} lattice_t;

static void lattice_init(lattice_t* L, int q, const long long* dims) {
  L->Q = q;
  L->D = q == 19 ? 3 : 2;
  L->n[0] = dims[0];
  L->n[1] = dims[1];
  L->n[2] = q == 19 ? dims[2] : 1;
  L->ncell = L->n[0] * L->n[1] * L->n[2];
  for (int k = 0; k < q; ++k)
    for (int d = 0; d < 3; ++d) L->e[k][d] = q == 19 ? E19[k][d] : E9[k][d];
  /* opp[k]: the direction with e = -e[k] (same search as lbm.py:360-380) */
  for (int k = 0; k < q; ++k)
    for (int p = 0; p < q; ++p)
      if (L->e[p][0] == -L->e[k][0] && L->e[p][1] == -L->e[k][1] && L->e[p][2] == -L->e[k][2])
        L->opp[k] = p;
  for (int k = 0; k < q; ++k) {
    if (q == 19) L->w[k] = k == 0 ? 1.0 / 3.0 : (k < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
    else L->w[k] = k == 0 ? 4.0 / 9.0 : (k < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
  }
}

/* cell index of (i0, i1, i2) + s*e_k with periodic wrap (lbm.py:242-267
   rolls by +e, so the streamed value at x comes from x - e_k) */
static inline long long nbr(const lattice_t* L, long long i0, long long i1, long long i2, int k,
                            int s) {
  long long j0 = i0 + s * L->e[k][0], j1 = i1 + s * L->e[k][1], j2 = i2 + s * L->e[k][2];
  j0 = j0 < 0 ? j0 + L->n[0] : (j0 >= L->n[0] ? j0 - L->n[0] : j0);
  j1 = j1 < 0 ? j1 + L->n[1] : (j1 >= L->n[1] ? j1 - L->n[1] : j1);
  j2 = j2 < 0 ? j2 + L->n[2] : (j2 >= L->n[2] ? j2 - L->n[2] : j2);
  return (j0 * L->n[1] + j1) * L->n[2] + j2;
}

/* faces: 0 none, 1..6 = x_min, x_max, y_min, y_max, z_min, z_max (lbm.py:517-524) */
static int face_ok(const lattice_t* L, int face) {
  return face >= 0 && face <= 6 && (face == 0 || (face - 1) / 2 < L->D);
}

#define REAL double
#define SFX(name) name##_f64
#include "lbm_ref.c"
#undef REAL
#undef SFX
#define REAL float
#define SFX(name) name##_f32
#include "lbm_ref.c"
#undef REAL
#undef SFX

/* The wrapper's check that it loaded the library this file describes. */
int lbm_oracle_version(void) { return 1; }

#else /* ---- body instantiated for REAL = double and REAL = float ---------- */

/* moments of one cell's populations with the F/2 correction (lbm.py:209-239) */
static inline void SFX(cell_moments)(const lattice_t* L, const REAL* fc, const REAL* F,
                                     REAL* rho, REAL u[3]) {
  REAL r = 0, m[3] = {0, 0, 0};
  for (int k = 0; k < L->Q; ++k) {
    r += fc[k];
    for (int d = 0; d < L->D; ++d) m[d] += fc[k] * (REAL)L->e[k][d];
  }
  const REAL den = r > (REAL)1e-10 ? r : (REAL)1e-10;
  for (int d = 0; d < 3; ++d) u[d] = d < L->D ? (m[d] + (REAL)0.5 * F[d]) / den : (REAL)0;
  *rho = r;
}

/* feq_k (lbm.py:180-206): w rho (1 + e.u/cs2 + (e.u)^2/(2 cs2^2) - u.u/(2 cs2)) */
static inline REAL SFX(feq)(const lattice_t* L, int k, REAL rho, const REAL u[3]) {
  const REAL cs2 = (REAL)(1.0 / 3.0);
  REAL eu = 0, usq = 0;
  for (int d = 0; d < L->D; ++d) {
    eu += u[d] * (REAL)L->e[k][d];
    usq += u[d] * u[d];
  }
  return (REAL)L->w[k] * rho *
         ((REAL)1 + eu / cs2 + eu * eu / ((REAL)2 * cs2 * cs2) - usq / ((REAL)2 * cs2));
}

/* S_k (lbm.py:307-353): (1 - 1/(2 tau)) w sum_d [(e_d - u_d)/cs2 + (e.u) e_d/cs2^2] F_d */
static inline REAL SFX(guo)(const lattice_t* L, int k, REAL tau, const REAL u[3], const REAL* F) {
  const REAL cs2 = (REAL)(1.0 / 3.0);
  REAL eu = 0;
  for (int d = 0; d < L->D; ++d) eu += u[d] * (REAL)L->e[k][d];
  REAL s = 0;
  for (int d = 0; d < L->D; ++d)
    s += (((REAL)L->e[k][d] - u[d]) / cs2 + (REAL)L->e[k][d] * (eu / (cs2 * cs2))) * F[d];
  return ((REAL)1 - (REAL)0.5 / tau) * (REAL)L->w[k] * s;
}

static void SFX(zou_he)(const lattice_t* L, REAL* f, const uint8_t* mask, int face, REAL rho_p) {
  const int axis = (face - 1) / 2, side_min = ((face - 1) % 2) == 0;
  int known[19], unknown[19], tang[19], nk = 0, nu = 0, nt = 0;
  /* lbm.py:383-422: unknown = pointing into the domain from this face */
  for (int k = 0; k < L->Q; ++k) {
    const int c = L->e[k][axis];
    if (c == 0) tang[nt++] = k;
    else if ((c > 0) == side_min) unknown[nu++] = k;
    else known[nk++] = k;
  }
  const long long fixed = side_min ? 0 : L->n[axis] - 1;
  const REAL r = rho_p > (REAL)1e-10 ? rho_p : (REAL)1e-10;
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < L->ncell; ++c) {
    long long i[3] = {c / (L->n[2] * L->n[1]), (c / L->n[2]) % L->n[1], c % L->n[2]};
    if (i[axis] != fixed || (mask && mask[c])) continue;
    REAL* fc = f + c * L->Q;
    REAL sk = 0, st = 0;
    for (int j = 0; j < nk; ++j) sk += fc[known[j]];
    for (int j = 0; j < nt; ++j) st += fc[tang[j]];
    const REAL s = sk + st;
    /* the reference's normal velocity, without a factor 2 on the known sum
       (lbm.py:476-479) */
    REAL u[3] = {0, 0, 0};
    u[axis] = side_min ? (REAL)1 - s / r : (REAL)-1 + s / r;
    for (int j = 0; j < nu; ++j) {
      const int k = unknown[j], o = L->opp[k];
      fc[k] = fc[o] + SFX(feq)(L, k, rho_p, u) - SFX(feq)(L, o, rho_p, u);
    }
  }
}

/* Collision of every cell. Q, D and the velocity table are passed as
   constants by the caller so that the per-cell loops unroll; the arithmetic
   is that of cell_moments / feq / guo above, term for term. */
static inline __attribute__((always_inline)) void SFX(collide_all)(
    const lattice_t* L, const int Q, const int D, const int (*E)[3], const REAL* f, REAL* post,
    const REAL* force, int force_field, REAL tau, const REAL* tau_field) {
  const REAL cs2 = (REAL)(1.0 / 3.0);
  const REAL zero3[3] = {0, 0, 0};
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < L->ncell; ++c) {
    const REAL* fc = f + c * Q;
    const REAL* F = !force ? zero3 : (force_field ? force + c * D : force);
    const REAL t = tau_field ? tau_field[c] : tau;
    REAL r = 0, m[3] = {0, 0, 0};
    for (int k = 0; k < Q; ++k) {
      r += fc[k];
      for (int d = 0; d < D; ++d) m[d] += fc[k] * (REAL)E[k][d];
    }
    const REAL den = r > (REAL)1e-10 ? r : (REAL)1e-10;
    REAL u[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) u[d] = (m[d] + (REAL)0.5 * F[d]) / den;
    REAL usq = 0;
    for (int d = 0; d < D; ++d) usq += u[d] * u[d];
    const REAL pref = (REAL)1 - (REAL)0.5 / t;
    for (int k = 0; k < Q; ++k) {
      REAL eu = 0;
      for (int d = 0; d < D; ++d) eu += u[d] * (REAL)E[k][d];
      const REAL wk = (REAL)L->w[k];
      const REAL fe =
          wk * r * ((REAL)1 + eu / cs2 + eu * eu / ((REAL)2 * cs2 * cs2) - usq / ((REAL)2 * cs2));
      REAL sg = 0;
      for (int d = 0; d < D; ++d)
        sg += (((REAL)E[k][d] - u[d]) / cs2 + (REAL)E[k][d] * (eu / (cs2 * cs2))) * F[d];
      post[c * Q + k] = fc[k] - (fc[k] - fe) / t + pref * wk * sg;
    }
  }
}

/* One step: f (AoS) -> f_out, with `post` as scratch of the same size.
 * force: NULL, D uniform values (force_field = 0) or (*G, D) (force_field = 1).
 * mask: NULL or (*G) uint8 (nonzero = wall; 2 = moving wall for bb = 2).
 * bb: 1 = reference node swap, 2 = halfway with lid velocity (extension).
 * rho, u, p: output moments (each nullable). */
int SFX(lbm_oracle_update)(int q, const long long* dims, const REAL* f, REAL* f_out, REAL* post,
                           const uint8_t* mask, int bb, const double* lid, const REAL* force,
                           int force_field, double tau, const REAL* tau_field, int inlet_face,
                           double inlet_rho, int outlet_face, double outlet_rho, REAL* rho,
                           REAL* u, REAL* p, int nthreads) {
  if ((q != 19 && q != 9) || !dims || !f || !f_out || !post || f == f_out) return -1;
  if (bb != 1 && bb != 2) return -1;
  lattice_t L;
  lattice_init(&L, q, dims);
  if (!face_ok(&L, inlet_face) || !face_ok(&L, outlet_face)) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int Q = L.Q, D = L.D;
  const REAL zero3[3] = {0, 0, 0};
  const REAL cs2 = (REAL)(1.0 / 3.0);
  /* 1-2: moments and collision on every cell, walls included (lbm.py:934-939) */
  if (Q == 19)
    SFX(collide_all)(&L, 19, 3, E19, f, post, force, force_field, (REAL)tau, tau_field);
  else
    SFX(collide_all)(&L, 9, 2, E9, f, post, force, force_field, (REAL)tau, tau_field);
  /* 3-4: streaming and the wall rule */
#pragma omp parallel for collapse(2) schedule(static)
  for (long long i0 = 0; i0 < L.n[0]; ++i0)
    for (long long i1 = 0; i1 < L.n[1]; ++i1)
      for (long long i2 = 0; i2 < L.n[2]; ++i2) {
        const long long c = (i0 * L.n[1] + i1) * L.n[2] + i2;
        const int m = mask ? mask[c] : 0;
        REAL* oc = f_out + c * Q;
        if (bb == 1) {
          /* streamed slot k at c came from c - e_k; a wall cell takes slot
             opp(k), which came from c - e_opp(k) = c + e_k (lbm.py:944-948) */
          for (int k = 0; k < Q; ++k)
            oc[k] = m ? post[nbr(&L, i0, i1, i2, k, +1) * Q + L.opp[k]]
                      : post[nbr(&L, i0, i1, i2, k, -1) * Q + k];
        } else if (m) {
          for (int k = 0; k < Q; ++k) oc[k] = (REAL)L.w[k];  /* solids parked at rest */
        } else {
          for (int k = 0; k < Q; ++k) {
            const long long src = nbr(&L, i0, i1, i2, k, -1);
            const int ms = (k == 0 || !mask) ? 0 : mask[src];
            if (ms) {
              REAL lidterm = 0;
              if (ms == 2) {
                double eu = 0;
                for (int d = 0; d < D; ++d) eu += L.e[k][d] * lid[d];
                lidterm = (REAL)(2.0 * L.w[k] * eu / (1.0 / 3.0));
              }
              oc[k] = post[c * Q + L.opp[k]] + lidterm;
            } else {
              oc[k] = post[src * Q + k];
            }
          }
        }
      }
  /* 5: Zou-He faces, inlet then outlet (lbm.py:950-968) */
  if (inlet_face) SFX(zou_he)(&L, f_out, mask, inlet_face, (REAL)inlet_rho);
  if (outlet_face) SFX(zou_he)(&L, f_out, mask, outlet_face, (REAL)outlet_rho);
  /* 6: output moments (lbm.py:970-976) */
  if (rho || u || p) {
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < L.ncell; ++c) {
      const REAL* F = !force ? zero3 : (force_field ? force + c * D : force);
      REAL r, uu[3];
      SFX(cell_moments)(&L, f_out + c * Q, F, &r, uu);
      const int wall = mask && mask[c];
      if (rho) rho[c] = r;
      if (u)
        for (int d = 0; d < D; ++d) u[c * D + d] = wall ? (REAL)0 : uu[d];
      if (p) p[c] = r * cs2;
    }
  }
  return 0;
}

/* nsteps updates, ping-ponging f_a / f_b; the result is in f_a when nsteps
 * is even and in f_b when it is odd; moments of the last step go to rho/u/p. */
int SFX(lbm_oracle_run)(int q, const long long* dims, REAL* f_a, REAL* f_b, REAL* post,
                        const uint8_t* mask, int bb, const double* lid, const REAL* force,
                        int force_field, double tau, const REAL* tau_field, int inlet_face,
                        double inlet_rho, int outlet_face, double outlet_rho, REAL* rho, REAL* u,
                        REAL* p, int nthreads, int nsteps) {
  for (int s = 0; s < nsteps; ++s) {
    REAL* src = (s % 2 == 0) ? f_a : f_b;
    REAL* dst = (s % 2 == 0) ? f_b : f_a;
    const int last = s == nsteps - 1;
    const int rc = SFX(lbm_oracle_update)(q, dims, src, dst, post, mask, bb, lid, force,
                                          force_field, tau, tau_field, inlet_face, inlet_rho,
                                          outlet_face, outlet_rho, last ? rho : NULL,
                                          last ? u : NULL, last ? p : NULL, nthreads);
    if (rc) return rc;
  }
  return 0;
}

#endif
