/*
 * lbm_c_api.h - C ABI of the MI355X (gfx950) lattice-Boltzmann kernels.
 *
 * One header for the whole library (libmaddening_lbm.so). Plain pointers and
 * sizes only: every pointer argument named f_*, mask, force_field, tau_field,
 * rho, u, p, work, *_bar is a DEVICE pointer; `stream` is a hipStream_t passed
 * as void*. No entry point allocates device memory: the caller passes every
 * buffer, including the workspace sized by lbm_workspace_size(). All work is
 * enqueued on `stream` and no entry point synchronises the host.
 *
 * Population layout ("SoA"): f[q][i0][i1][i2] for D3Q19 and f[q][i0][i1] for
 * D2Q9, i.e. the reference's (*grid_shape, Q) array transposed to
 * (Q, *grid_shape), last grid axis contiguous. Direction order, weights and
 * opposite map are those of the reference's d3q19()/d2q9()
 * (src/maddening/nodes/lbm.py:90-173).
 *
 * Storage types: LBM_F32/LBM_F64 hold f itself; LBM_F32_DEV, LBM_F16_DEV and
 * LBM_BF16_DEV hold the deviation f_q - w_q (LBM_F16_DEV scaled by 2^10 so
 * that small deviations stay out of the subnormal range). Collision arithmetic is FP32
 * for every storage type but F64, which computes in FP64.
 *
 * Auxiliary per-cell arrays use the reference's layouts (last grid axis
 * contiguous): mask uint8 (*G) (0 fluid, 1 wall, 2 moving wall),
 * force_field float (*G, D), tau_field float (*G), rho float (*G),
 * u float (*G, D), p float (*G).
 *
 * Return value of every int function: 0 on success, a negative lbm_status on
 * failure; lbm_last_error() then gives a message (thread-local).
 */
#ifndef MADDENING_LBM_C_API_H
#define MADDENING_LBM_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_API_VERSION 8

typedef enum lbm_status {
  LBM_OK = 0,
  LBM_ERR_INVALID = -1,     /* bad argument (null pointer, bad size, tau <= 0.5 ...) */
  LBM_ERR_UNSUPPORTED = -2, /* combination not implemented */
  LBM_ERR_HIP = -3          /* a HIP runtime call failed */
} lbm_status;

typedef enum lbm_storage {
  LBM_F32 = 0,      /* float f */
  LBM_F64 = 1,      /* double f */
  LBM_F16_DEV = 2,  /* IEEE half, (f - w) * 1024 */
  LBM_BF16_DEV = 3, /* bfloat16, f - w */
  LBM_F32_DEV = 4   /* float, f - w */
} lbm_storage;

typedef enum lbm_bb_mode {
  LBM_BB_NONE = 0,      /* fully periodic, mask ignored */
  LBM_BB_REFERENCE = 1, /* the reference's node swap: f_q(x) <- f_stream_opp(q)(x) at wall cells */
  LBM_BB_HALFWAY = 2    /* link-wise halfway bounce-back, moving wall for mask value 2 */
} lbm_bb_mode;

typedef enum lbm_scheme {
  LBM_SCHEME_AUTO = 0,
  LBM_SCHEME_PULL = 1,  /* stored state between kernels is post-collision */
  LBM_SCHEME_PUSH = 2   /* stored state between kernels is pre-collision */
} lbm_scheme;

/* Lattice description shared by every entry point. */
typedef struct lbm_desc {
  int32_t q;            /* 9 (D2Q9) or 19 (D3Q19) */
  int32_t storage;      /* lbm_storage */
  int32_t bb_mode;      /* lbm_bb_mode */
  int32_t scheme;       /* lbm_scheme; AUTO picks the measured-fastest */
  int64_t dims[3];      /* grid shape in the reference's axis order; dims[2] ignored for D2Q9 */
  double tau;           /* BGK relaxation time used where tau_field is NULL; must be > 0.5 */
  double force[3];      /* uniform body force used where force_field is NULL */
  double lid_velocity[3]; /* velocity of mask==2 cells in LBM_BB_HALFWAY */
  int32_t block_x;      /* threads per workgroup of the one-cell-per-thread kernels (64, 128,
                           192 or 256); -1 = the pull step over 256-cell tiles of the
                           flattened grid; 0 = auto (row blocks of 256 when the contiguous
                           axis is a multiple of 256 cells, flattened tiles otherwise) */
  int32_t row_vec;      /* periodic lattices only: cells per lane of the row-per-wave kernels
                           (1, 2 or 4; the contiguous axis must be a multiple of 64*row_vec);
                           0 = auto, -1 = never */
  int32_t flags;        /* LBM_FLAG_* */
  int32_t xcd_group;    /* with LBM_FLAG_XCD_REMAP: row-blocks per run dealt to one XCD
                           (1 = identity order, 0 = one contiguous eighth of the lattice
                           per XCD); without the flag the library picks: runs of 512 when
                           the lattice has at most 24576 row-blocks and dims[1]*dims[2]
                           cells per a-plane (D2Q9: dims[0]*dims[1]) is a power of two,
                           identity otherwise */
  /* Zou-He pressure faces (lbm.py:425-524, 950-968), applied after streaming,
   * inlet first: 0 = none, 1..6 = x_min, x_max, y_min, y_max, z_min, z_max
   * in the reference's axis order (_FACE_MAP, lbm.py:517-524). */
  int32_t inlet_face;
  int32_t outlet_face;
  double inlet_rho;     /* prescribed density = inlet_pressure / cs2 */
  double outlet_rho;
} lbm_desc;

/* flags */
#define LBM_FLAG_XCD_REMAP 1  /* use xcd_group as given instead of the library's choice */
/* Adjoint cotangent convention under LBM_BB_REFERENCE (ignored for
 * LBM_BB_NONE, rejected for LBM_BB_HALFWAY): the cotangent buffer holds the
 * slots q and opp(q) of every wall cell exchanged. In that form the transpose
 * of the node-swap bounce-back is a plain stencil gather, so a chain of
 * lbm_adjoint_step calls reads no neighbour mask. IN applies to fbar_next /
 * fbar_in, OUT to fbar_out. */
#define LBM_FLAG_ADJ_IN_SWAPPED 2
#define LBM_FLAG_ADJ_OUT_SWAPPED 4
/* Adjoint with Zou-He faces: fbar_out is passed back through the transpose
 * of the faces that produced the step's starting state (that is, it is the
 * cotangent with respect to the state before those faces) and dL/drho of
 * each face is added to work. Set it on every lbm_adjoint_step of a chain
 * except the one that returns dL/df_0, and on lbm_moments_adjoint. */
#define LBM_FLAG_ADJ_OUT_FACES 8
/* Two steps per launch (no counterpart in the reference: a launch schedule
 * of the same update). lbm_step runs its pull sequence two units per launch
 * (a collision, the fused steps, the final stream + moments), keeping the
 * intermediate state in LDS, where the lattice allows it: D3Q19 with
 * LBM_F32 or LBM_F32_DEV, LBM_BB_NONE or LBM_BB_REFERENCE, with or without
 * force and tau fields and pressure faces, and D3Q19 with LBM_F16_DEV or
 * LBM_BF16_DEV, periodic without force or fields (under PAIR_FORCE also with
 * them); D2Q9 periodic with uniform force and tau, its fused steps only
 * (under PAIR_FORCE also LBM_F16_DEV and LBM_BF16_DEV); Q * cells < 2^31.
 * By default only where it measured faster (D3Q19: at least 2^21 cells,
 * dims[0] and dims[1] at least 64, dims[2] one of 64, 96, ..., 256 or above
 * 341; D2Q9: at least 2^21 cells); PAIR_FORCE uses it wherever it
 * applies, PAIR_OFF never. A single D3Q19 update
 * (nsteps = 1; also 16-bit storage) is one launch unless PAIR_OFF. Results
 * are bit-identical either way. */
#define LBM_FLAG_PAIR_OFF 16
#define LBM_FLAG_PAIR_FORCE 32
/* Three units in the first launch (no counterpart in the reference): a
 * D3Q19 sequence run two units per launch has an odd number of units when
 * nsteps is even (4 or more), and lbm_step then runs the collision and the
 * first two fused steps as one three-unit launch instead of a two-unit
 * launch and a single one, where it applies: 32- or 16-bit storage (not
 * LBM_F64), LBM_BB_NONE, no force and no fields, dims[2] a multiple of 64,
 * the state within 4 GiB (19 * dims product * 4 or 2 bytes and a plane
 * more). Bit-identical; TRI_OFF turns it off. */
#define LBM_FLAG_TRI_OFF 64

/* ---- queries ----------------------------------------------------------- */

/* Library ABI version (LBM_API_VERSION). */
int lbm_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char* lbm_last_error(void);

/* Bytes of device workspace lbm_step() needs for `nsteps` steps (0 if none):
 * one state for the pull sequence's ping-pong, plus 8 bytes per cell of
 * bounce-back link words under LBM_BB_HALFWAY. */
size_t lbm_workspace_size(const lbm_desc* d, int nsteps);

/* Bytes of one population state (Q * cells * storage size). */
size_t lbm_state_bytes(const lbm_desc* d);

/* ---- forward ----------------------------------------------------------- */

/*
 * Advance `nsteps` (>= 1) steps of LBMNode.update
 * (src/maddening/nodes/lbm.py:899-986): Guo-forced BGK collision on every
 * cell, periodic streaming (lbm.py:242-267), bounce-back per d->bb_mode
 * (reference mode: lbm.py:944-948), then the output moments of the final
 * state (lbm.py:970-976): rho, u with the F/2 correction and u = 0 at mask
 * cells, p = rho/3. rho/u/p may each be NULL to skip them.
 * f_in and f_out are SoA states in d->storage and must not overlap.
 * mask (nullable) selects wall cells: it plays the role of the reference's
 * wall_mask_update / state["wall_mask"] (lbm.py:914-919).
 * force_field (nullable) replaces d->force per cell (body_force of shape
 * (*G, D), lbm.py:922-927); tau_field (nullable) replaces d->tau per cell.
 * work: lbm_workspace_size(d, nsteps) bytes (may be NULL when that is 0).
 */
int lbm_step(void* stream, const lbm_desc* d, const void* f_in, void* f_out,
             void* work, const uint8_t* mask, const float* force_field,
             const float* tau_field, float* rho, float* u, float* p,
             int nsteps);

/*
 * One launch of the pull sequence lbm_step runs (its building block, for
 * callers that keep post-collision states, like the adjoint's recompute):
 *   phase 0: src = a state f_n, dst = its collision C(f_n);
 *   phase 1: src = post-collision g_{n-1}, dst = C(f_n) with f_n gathered
 *            from g_{n-1} by streaming and bounce-back;
 *   phase 2: src = post-collision g_{n-1}, dst = f_n, and the moments of f_n
 *            into rho/u/p (each nullable);
 *   phase 3: src = post-collision g_{n-1}, dst = C(f_{n+1}): two phase-1
 *            launches in one, bit for bit (LBM_FLAG_PAIR_*); returns
 *            LBM_ERR_UNSUPPORTED where the two-step launch does not apply;
 *   phases 4, 5, 6: the other D3Q19 two-unit launches, as phase 0 then 1,
 *            phase 1 then 2 (rho/u/p as in phase 2), and phase 0 then 2;
 *   phase 7: src = a state f_n, dst = C(f_{n+2}): phase 0 and two phase-1
 *            launches in one, bit for bit, where LBM_FLAG_TRI_OFF's rule
 *            allows it but for the row length (any dims[2] >= 64);
 *            LBM_ERR_UNSUPPORTED elsewhere.
 * links (nullable, LBM_BB_HALFWAY only): the link words lbm_links_build made
 * from this mask; the gathers then read one word per cell instead of the
 * neighbours' mask bytes (the same results).
 */
int lbm_pull_phase(void* stream, const lbm_desc* d, int phase, const void* src, void* dst,
                   const uint8_t* mask, const uint32_t* links, const float* force_field,
                   const float* tau_field, float* rho, float* u, float* p);

/*
 * Launches lbm_step makes for nsteps steps with this descriptor (with_fields:
 * the call passes a force or tau field), and how many of them run two or
 * three units (LBM_FLAG_PAIR_*, LBM_FLAG_TRI_OFF). Pull sequence: collide,
 * two-step launches, single fused launches, stream + moments. Push scheme: one launch per step
 * plus the moments pass (Zou-He faces add one launch per face and step).
 * No counterpart in the reference; for timing and profiling.
 */
int lbm_step_schedule(const lbm_desc* d, int nsteps, int with_fields, int* launches,
                      int* pair_launches);

/*
 * Link words of a wall mask for LBM_BB_HALFWAY (no counterpart in the
 * reference: they restate its mask, lbm.py:914-919, per link). Per cell, two
 * 32-bit words: which directions q have a wall at x - e_q, whether one of
 * those walls moves (mask 2) and whether the cell is a wall; then which links
 * end at a moving wall. lbm_links_size bytes of device memory; rebuild after
 * the mask changes. lbm_step builds its own inside its workspace.
 */
size_t lbm_links_size(const lbm_desc* d);
int lbm_links_build(void* stream, const lbm_desc* d, const uint8_t* mask, uint32_t* links);

/*
 * Output moments of a state without stepping: _compute_macroscopic with the
 * Guo F/2 half-correction (lbm.py:209-239), u zeroed at mask cells and
 * p = rho * cs2 (lbm.py:971-976).
 */
int lbm_moments(void* stream, const lbm_desc* d, const void* f,
                const uint8_t* mask, const float* force_field, float* rho,
                float* u, float* p);

/*
 * f = feq(rho, u) (lbm.py:180-206) into a state in d->storage.
 * rho (*G) and u (*G, D) float; rho may be NULL (rho = 1), u may be NULL (u = 0).
 */
int lbm_equilibrium(void* stream, const lbm_desc* d, const float* rho,
                    const float* u, void* f_out);

/*
 * Layout/dtype conversion between the reference's (*G, Q) array
 * (lbm.py:18-19) and the SoA state. to_soa != 0: aos -> soa; else soa -> aos.
 * aos_is_f64 selects float64 (*G, Q) instead of float32 (*G, Q).
 */
int lbm_convert(void* stream, const lbm_desc* d, int to_soa, int aos_is_f64,
                const void* src, void* dst);

/*
 * FP64 conservation sums of a state: out[0] = sum over cells and q of f,
 * out[1..D] = sum of f * e_q (raw first moment, no force correction).
 * out: device double[4]. work: lbm_reduce_workspace_size(d) bytes.
 */
size_t lbm_reduce_workspace_size(const lbm_desc* d);
int lbm_conservation(void* stream, const lbm_desc* d, const void* f,
                     void* work, double* out);

/*
 * lbm_conservation with no workspace argument and no allocation: its
 * per-workgroup partial sums live in a 32 KiB array the library holds on
 * each GPU it runs on, until the process ends. No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It calls no hipMalloc, hipFree or allocating routine. The
 * first call on a device, made with that device current, looks up the
 * array's address. Every call on one GPU shares its array: calls that
 * overlap on one GPU (on different streams) race, so order them on one
 * stream. Same results as lbm_conservation, bit for bit.
 */
int lbm_conservation_static(void* stream, const lbm_desc* d, const void* f, double* out);

/* ---- batches of independent lattices ------------------------------------ */
/*
 * No counterpart in the reference: its only batching is jax.vmap over
 * equal-shaped states in GraphManager.run_sweep (graph_manager.py:2807-2877).
 * Here lattices of any size, lattice type, tau, force and wall rule advance
 * together (float32 storage): each step is one launch per lattice class, and
 * a whole run can be replayed as one HIP graph.
 */
typedef struct lbm_batch_item {
  int32_t q;              /* 9 or 19 */
  int32_t bb_mode;        /* lbm_bb_mode */
  int64_t dims[3];        /* grid shape, reference axis order (dims[2] ignored for D2Q9) */
  double tau;
  double force[3];        /* uniform body force */
  double lid_velocity[3]; /* mask==2 cells under LBM_BB_HALFWAY */
  void* f[2];             /* two float32 SoA state buffers; a run starts from f[0] */
  const uint8_t* mask;    /* required unless bb_mode == LBM_BB_NONE */
  float* rho;             /* nullable moment outputs written by the last launch of a run */
  float* u;
  float* p;
  int64_t pitch;          /* cells from the start of one population plane of f[0]/f[1] to the
                             next: 0 = the cell count (the lbm_step layout); else at least the
                             cell count. A multiple of 64 keeps every store of a step aligned */
  uint32_t* links;        /* LBM_BB_HALFWAY only, nullable: 8 bytes per cell of device memory
                             for link words, rebuilt from the mask at the start of every run;
                             the steps then read one word per cell instead of the neighbours'
                             mask bytes (the same results) */
} lbm_batch_item;

/* Opaque host-side plan: the class ranges the launches need. */
typedef struct lbm_batch_plan_t lbm_batch_plan_t;

/* Bytes of device memory for the plan table of `nitems` lattices. */
size_t lbm_batch_plan_size(int nitems);

/*
 * Build a plan: validates every item, orders the items into classes of equal
 * (q, bb_mode) (periodic D3Q19 lattices with rows of at most 341 cells form
 * a class of their own, see lbm_batch_plan_pairs), cuts each lattice into
 * tiles of 256 * cells_per_thread
 * consecutive cells (1..16; 0 = 1) and uploads the table to `device_plan`
 * (lbm_batch_plan_size bytes, owned by the caller, alive as long as the
 * plan). A set-up call: it waits for the upload on `stream`. *total_tiles
 * (nullable) receives the tile count. Free the handle with
 * lbm_batch_plan_destroy.
 */
int lbm_batch_plan_create(void* stream, const lbm_batch_item* items, int nitems,
                          int cells_per_thread, void* device_plan, lbm_batch_plan_t** plan,
                          long long* total_tiles);

/*
 * Two steps per launch for the periodic D3Q19 lattices of a plan (rows of
 * at most 341 cells): on (the default after lbm_batch_plan_create: the
 * 64-lattice benchmark batch, whose 32 periodic cubes have edges of 32-160
 * cells, runs at 44.8k MLUPS against 40.5k with one step per launch), or off.
 * Their fused steps then run two per launch, in an even number of launches,
 * so every lattice ends in the same buffer with the same bits as with one
 * launch per step. Takes effect for later lbm_batch_run and
 * lbm_batch_graph_create calls. No counterpart in the reference.
 */
int lbm_batch_plan_pairs(lbm_batch_plan_t* plan, int on);

int lbm_batch_plan_destroy(lbm_batch_plan_t* plan);

/*
 * nsteps LBMNode.update steps of every lattice in the plan: per step one
 * launch per lattice class present (at most 6), each striding `grid`
 * workgroups over that class's tiles (nsteps + 1 rounds: the pull sequence).
 * The final state of each item is in f[(nsteps + 1) % 2].
 */
int lbm_batch_run(void* stream, const lbm_batch_plan_t* plan, int nsteps, int grid);

/* The same run captured once into a HIP graph; launch it as often as needed. */
int lbm_batch_graph_create(const lbm_batch_plan_t* plan, int nsteps, int grid, void** graph);
int lbm_batch_graph_launch(void* stream, void* graph);
int lbm_batch_graph_destroy(void* graph);

/* ---- adjoint (reverse mode) -------------------------------------------- */
/*
 * No counterpart in the reference, which differentiates lbm.py:899-986 with
 * jax.grad through lax.scan (tests/verification/test_gradient_health.py:56-75
 * shows the pattern). These entry points are the hand-written transpose of
 * one LBMNode.update, for f32 or f64 storage. Cotangent buffers (fbar_*) are
 * SoA like the state and hold plain values in float (f32 storage) or double
 * (f64 storage).
 *
 * Per-workgroup partial sums of dL/dF for a uniform force live in `work`
 * (lbm_adjoint_workspace_size bytes): lbm_adjoint_begin zeroes them, every
 * lbm_adjoint_step / lbm_moments_adjoint adds to them, lbm_adjoint_end folds
 * them into force_bar_out (device double[3], reference axis order).
 */
size_t lbm_adjoint_workspace_size(const lbm_desc* d);
int lbm_adjoint_begin(void* stream, const lbm_desc* d, void* work);

/*
 * fbar_out = (dStep/df at f_n)^T fbar_next for one step that started from
 * f_n. With f_n_post == 0, `f_n` is that state; with f_n_post != 0 it is the
 * post-collision state of the previous step (what lbm_pull_phase phases 0
 * and 1 write) and f_n is gathered from it by the forward's own streaming
 * and bounce-back. Adds dL/dtau(x) into tau_bar (*G) and dL/dF(x) into
 * force_bar_field (*G, D) when those are non-NULL; adds uniform-force
 * partials into work when work is non-NULL. links: as in lbm_pull_phase.
 * The other arguments mean what they mean in lbm_step.
 */
int lbm_adjoint_step(void* stream, const lbm_desc* d, const void* f_n, int f_n_post,
                     const void* fbar_next, void* fbar_out, const uint8_t* mask,
                     const uint32_t* links, const float* force_field, const float* tau_field,
                     float* tau_bar, float* force_bar_field, void* work);

/*
 * Pull the cotangents of the output moments of state f (rho_bar, u_bar,
 * p_bar; each nullable) back onto f: fbar_out = fbar_in (nullable) +
 * d(moments)/df^T (...), following the moments of lbm.py:970-976 (F/2
 * correction, u = 0 at mask cells, p = rho/3). dL/dF from the F/2 term goes
 * to force_bar_field / work as in lbm_adjoint_step. fbar_in may equal fbar_out.
 */
int lbm_moments_adjoint(void* stream, const lbm_desc* d, const void* f,
                        const uint8_t* mask, const float* force_field,
                        const float* rho_bar, const float* u_bar,
                        const float* p_bar, const void* fbar_in, void* fbar_out,
                        float* force_bar_field, void* work);

int lbm_adjoint_end(void* stream, const lbm_desc* d, const void* work,
                    double* force_bar_out);

/* dL/drho_in, dL/drho_out of the Zou-He faces (prescribed densities, i.e.
 * pressure / cs2) accumulated in work by calls with LBM_FLAG_ADJ_OUT_FACES:
 * rho_bar_out[0] inlet, [1] outlet (0 where a face is not configured). */
int lbm_adjoint_face_end(void* stream, const lbm_desc* d, const void* work,
                         double* rho_bar_out);

/* ---- pipe node ----------------------------------------------------------- */
/*
 * LBMPipeNode.update (src/maddening/nodes/lbm_pipe.py:751-879): D3Q19 fluid
 * in a pipe driven by an actuator-disc propeller and gravity. With
 * multiphase == 0 it carries a D3Q7 passive tracer (BGK, tau_tracer,
 * collided with the step's output velocity). With multiphase != 0 it adds a
 * Shan-Chen pseudopotential force with exact-difference forcing, the 0.25
 * velocity clamp and the 0.01 density floor; the tracer there is derived
 * from density and tracer_f passes through.
 *
 * Unlike the main path, velocities are plain sum(f e)/rho, without the F/2
 * correction. State: SoA f[19][nx][ny][nz] and g[7][nx][ny][nz], plain
 * values in f32 or f64 storage. cell_flags[(*G)]: bit 0 wall, bit 1
 * propeller disc. rho (*G), u (*G, 3) and tracer (*G) are written in the
 * storage's element type, and each is nullable. work is
 * lbm_pipe_workspace_size bytes.
 */
typedef struct lbm_pipe_desc {
  int64_t dims[3];         /* nx, ny, nz */
  int32_t storage;         /* LBM_F32 or LBM_F64 */
  int32_t multiphase;      /* 0: single phase + tracer; 1: Shan-Chen */
  double tau, tau_tracer;
  double propeller_strength; /* x-force on propeller-disc cells */
  double gravity;          /* z-force on fluid cells */
  double G, rho_0, rho_wall, rho_liquid, rho_gas;
  int32_t flags, reserved;
} lbm_pipe_desc;

size_t lbm_pipe_workspace_size(const lbm_pipe_desc* d);
int lbm_pipe_step(void* stream, const lbm_pipe_desc* d, const void* f_in, void* f_out,
                  const void* g_in, void* g_out, const uint8_t* cell_flags, void* work,
                  void* rho, void* u, void* tracer, int nsteps);

#ifdef __cplusplus
}
#endif

#endif /* MADDENING_LBM_C_API_H */
