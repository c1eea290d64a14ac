// Many independent lattices (D2Q9 and D3Q19, any sizes, own tau / force /
// walls) advanced by ONE launch per step.
//
// The plan is a table in device memory: one entry per lattice and a running
// tile count. A tile is 256 * cells_per_thread consecutive cells of one
// lattice (row-major over the whole grid, so short rows leave no thread
// idle); every lattice is cut into equal-sized pieces whatever its shape.
// A persistent grid of workgroups strides over all tiles of all lattices; a
// workgroup finds its lattice by binary search over the tile offsets and runs
// the same per-cell update as the single-lattice kernels (lbm_cell.hpp).
// Steps follow the pull sequence: launch k reads f[k % 2] and writes
// f[(k + 1) % 2]; launch 0 collides, launches 1 .. n-1 stream and collide,
// launch n streams (and writes moments where an entry asks for them).
// A whole n-step run can be captured once into a HIP graph and replayed.
//
// Periodic D3Q19 lattices with rows of at most 341 cells form a class of
// their own, which runs its fused steps two per launch (k_batch_pair, the
// march of lbm_pair.hpp (its LDS-ring form) over a second tile table: one 1024-thread
// workgroup per b-tile and a-chunk of each lattice). The number of such
// launches is even, so every lattice's final state is in the buffer where
// one launch per step leaves it, and the results are the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm_c_api.h"
#include "lbm_cell.hpp"
#include "lbm_common.hpp"
#include "lbm_host.hpp"
#include "lbm_pair.hpp"

namespace mlbm {

// planes per a-chunk of a two-step workgroup in a batch (the 32 periodic
// cubes of edge 32-160: 2685 workgroups; the time model rounds x (chunk + 2)
// planes puts 16 and 32 within 2 % of each other, 64 13 % behind)
constexpr int kBatchPairChunk = 32;
constexpr int kPairXcdRun = 16;  // consecutive two-step tiles dealt to one XCD

struct BatchEntry {
  int q, bb;
  int na, nb, nc, bpr;
  long long plane;       // cells
  long long pitch;       // cells from one population plane to the next
  long long tile_begin;
  long long pair_begin;  // first two-step tile (two-step classes only)
  int tb, btiles, la;    // two-step tiling: output rows, b-tiles, a-chunk planes
  float inv_tau, Fa, Fb, Fc, lid_a, lid_b, lid_c;
  void* buf[2];
  const uint8_t* mask;
  const uint32_t* links;  // halfway link words, or null (mask path)
  float* rho;
  float* u;
  float* p;
};

struct BatchHeader {
  long long total_tiles;
  int nitems;
  int cells_per_thread;
};

constexpr size_t kHeaderBytes = 64;  // keeps the entry table 64-byte aligned

__host__ __device__ inline const BatchEntry* entries(const void* plan) {
  return reinterpret_cast<const BatchEntry*>(static_cast<const char*>(plan) + kHeaderBytes);
}

// PHASE 0: collide only; 1: stream + collide; 2: stream (+ moments).
// One launch covers the tiles [t0, t1) of the entries [e0, e1), which all
// share the lattice L and wall rule BB.
template <class L, int BB, int PHASE>
__global__ void __launch_bounds__(256)
    k_batch(const void* plan, int parity, long long t0, long long t1, int e0, int e1) {
  const BatchHeader* h = static_cast<const BatchHeader*>(plan);
  const BatchEntry* items = entries(plan);
  const int cpt = h->cells_per_thread;
  constexpr bool PULL = PHASE != 0;
  constexpr bool COLLIDE = PHASE != 2;
  for (long long t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
    int lo = e0, hi = e1 - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchEntry& e = items[lo];
    // per-lattice cell indices fit int32 (check_desc caps a plane below
    // 2^31), so the row split is a 32-bit division, not a 64-bit one
    const int base = (int)(t - e.tile_begin) * (256 * cpt) + (int)threadIdx.x;
    const int plane = (int)e.plane;
    FwdArgs<float> a = {};
    a.src = e.buf[parity];
    a.dst = e.buf[parity ^ 1];
    a.mask = e.mask;
    a.links = e.links;
    a.links_moving = e.links ? e.links + e.plane : nullptr;
    a.plane = e.pitch;  // cell_update uses plane as the stride between planes
    a.na = e.na; a.nb = e.nb; a.nc = e.nc;
    a.inv_tau = e.inv_tau;
    a.Fa = e.Fa; a.Fb = e.Fb; a.Fc = e.Fc;
    a.lid_a = e.lid_a; a.lid_b = e.lid_b; a.lid_c = e.lid_c;
    if (PHASE == 2) {
      a.rho = e.rho; a.u = e.u; a.p = e.p;
    }
    for (int k = 0; k < cpt; ++k) {
      const int cell = base + 256 * k;
      if (cell >= plane) break;
      const int row = (int)((unsigned)cell / (unsigned)e.nc);
      const int ic = cell - row * e.nc;
      cell_update<L, float, BB, PULL, COLLIDE, true>(a, row, ic);
    }
  }
}

// A lattice class: entries [e0, e1) with tiles [t0, t1), same (q, wall rule);
// `pair`: periodic D3Q19 entries that also have two-step tiles [p0, p1).
struct BatchGroup {
  int q, bb, pair;
  int e0, e1;
  long long t0, t1;
  long long p0, p1;
};

// Two fused steps of every lattice of a two-step class: workgroup t of
// [p0, p1) runs b-tile and a-chunk t - pair_begin of its entry.
__global__ void __launch_bounds__(1024)
    k_batch_pair(const void* plan, int parity, long long p0, int nblocks, int remap, int e0,
                 int e1) {
  __shared__ PairRing x;
  const BatchEntry* items = entries(plan);
  const long long t = p0 + xcd_remap(blockIdx.x, nblocks, remap);
  int lo = e0, hi = e1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].pair_begin <= t) lo = mid;
    else hi = mid - 1;
  }
  // every field is the same across the workgroup: read it once into scalar
  // registers, so the march keeps its vector registers (no spills)
  auto u32 = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  auto uf = [&](float v) { return __int_as_float(u32(__float_as_int(v))); };
  auto up = [&](void* v) {
    const unsigned long long w = (unsigned long long)v;
    const unsigned lo32 = (unsigned)u32((int)(w & 0xffffffffu));
    const unsigned hi32 = (unsigned)u32((int)(w >> 32));
    return (void*)(((unsigned long long)hi32 << 32) | lo32);
  };
  const BatchEntry& e = items[u32(lo)];
  const int local = u32((int)(t - e.pair_begin));
  PairArgs a;
  a.src = up(e.buf[parity]);
  a.dst = up(e.buf[parity ^ 1]);
  a.na = u32(e.na); a.nb = u32(e.nb); a.nc = u32(e.nc);
  a.pitch = u32((int)e.pitch);
  a.w = a.nc; a.tc = a.nc; a.halo = 0; a.ctiles = 1;  // whole rows
  a.tb = u32(e.tb); a.btiles = u32(e.btiles); a.la = u32(e.la);
  a.nblocks = 0; a.remap = 0;
  a.inv_tau = uf(e.inv_tau);
  a.Fa = uf(e.Fa); a.Fb = uf(e.Fb); a.Fc = uf(e.Fc);
  pair_march_ring<float>(a, local % a.btiles, local / a.btiles, x);
}

template <class L, int BB>
int group_launch(hipStream_t s, const void* dplan, const BatchGroup& gr, int phase, int parity,
                 int grid) {
  const long long tiles = gr.t1 - gr.t0;
  const unsigned g = (unsigned)(grid < tiles ? grid : tiles);
  switch (phase) {
    case 0:
      hipLaunchKernelGGL((k_batch<L, BB, 0>), dim3(g), dim3(256), 0, s, dplan, parity, gr.t0,
                         gr.t1, gr.e0, gr.e1);
      break;
    case 1:
      hipLaunchKernelGGL((k_batch<L, BB, 1>), dim3(g), dim3(256), 0, s, dplan, parity, gr.t0,
                         gr.t1, gr.e0, gr.e1);
      break;
    default:
      hipLaunchKernelGGL((k_batch<L, BB, 2>), dim3(g), dim3(256), 0, s, dplan, parity, gr.t0,
                         gr.t1, gr.e0, gr.e1);
  }
  return check_launch("k_batch launch");
}

}  // namespace mlbm

namespace mlbm {
// A halfway item whose link words are rebuilt at the start of each run.
struct LinkJob {
  int q;
  const uint8_t* mask;
  uint32_t* links;
  Geom g;
};
}  // namespace mlbm

// Host handle of a plan: the device table plus the class ranges the
// launches need (host memory; the device table is the caller's buffer).
struct lbm_batch_plan_t {
  const void* dplan;
  std::vector<mlbm::BatchGroup> groups;
  std::vector<mlbm::LinkJob> links;
  // run two-step classes two steps per launch (lbm_batch_plan_pairs); on by
  // default: the 64-lattice benchmark batch ran at 44.8k MLUPS against 40.5k
  int pairs = 1;
};

namespace mlbm {

// which: 0 every class, 1 classes stepped one step per launch, 2 two-step classes
inline int batch_launch(hipStream_t s, const lbm_batch_plan_t* p, int phase, int parity,
                        int grid, int which = 0) {
  for (const BatchGroup& gr : p->groups) {
    const bool two = p->pairs && gr.pair;
    if ((which == 1 && two) || (which == 2 && !two)) continue;
    int rc;
    if (gr.q == 19) {
      if (gr.bb == BB_NONE) rc = group_launch<Lattice<19>, BB_NONE>(s, p->dplan, gr, phase, parity, grid);
      else if (gr.bb == BB_REF) rc = group_launch<Lattice<19>, BB_REF>(s, p->dplan, gr, phase, parity, grid);
      else if (gr.bb == BB_HALF) rc = group_launch<Lattice<19>, BB_HALF>(s, p->dplan, gr, phase, parity, grid);
      else rc = group_launch<Lattice<19>, BB_HALF_LINKS>(s, p->dplan, gr, phase, parity, grid);
    } else {
      if (gr.bb == BB_NONE) rc = group_launch<Lattice<9>, BB_NONE>(s, p->dplan, gr, phase, parity, grid);
      else if (gr.bb == BB_REF) rc = group_launch<Lattice<9>, BB_REF>(s, p->dplan, gr, phase, parity, grid);
      else if (gr.bb == BB_HALF) rc = group_launch<Lattice<9>, BB_HALF>(s, p->dplan, gr, phase, parity, grid);
      else rc = group_launch<Lattice<9>, BB_HALF_LINKS>(s, p->dplan, gr, phase, parity, grid);
    }
    if (rc) return rc;
  }
  return LBM_OK;
}

inline int batch_run(hipStream_t s, const lbm_batch_plan_t* p, int nsteps, int grid) {
  for (const LinkJob& j : p->links) {
    if (j.q == 19)
      hipLaunchKernelGGL((k_links<Lattice<19>>), dim3(j.g.nblocks), dim3(j.g.bx), 0, s, j.mask,
                         j.links, j.links + j.g.plane, j.g.na, j.g.nb, j.g.nc, j.g.bpr);
    else
      hipLaunchKernelGGL((k_links<Lattice<9>>), dim3(j.g.nblocks), dim3(j.g.bx), 0, s, j.mask,
                         j.links, j.links + j.g.plane, j.g.na, j.g.nb, j.g.nc, j.g.bpr);
    if (int rc = check_launch("k_links launch")) return rc;
  }
  int rc = batch_launch(s, p, 0, 0, grid);
  bool two = false;
  for (const BatchGroup& gr : p->groups) two = two || (p->pairs && gr.pair);
  for (int k = 1; k < nsteps && rc == LBM_OK; ++k)
    rc = batch_launch(s, p, 1, k & 1, grid, two ? 1 : 0);
  if (two) {
    // an even number of two-step launches keeps each lattice's parity: after
    // them the state is back in buf[1], where step 1 + 2 * npair expects it
    const int npair = 2 * ((nsteps - 1) / 4);
    int par = 1;
    for (int i = 0; i < npair && rc == LBM_OK; ++i, par ^= 1) {
      for (const BatchGroup& gr : p->groups) {
        if (!gr.pair) continue;
        const long long n = gr.p1 - gr.p0;
        hipLaunchKernelGGL(k_batch_pair, dim3((unsigned)n), dim3(1024), 0, s, p->dplan, par,
                           gr.p0, (int)n, kPairXcdRun, gr.e0, gr.e1);
        if ((rc = check_launch("k_batch_pair launch"))) break;
      }
    }
    for (int k = 1 + 2 * npair; k < nsteps && rc == LBM_OK; ++k)
      rc = batch_launch(s, p, 1, k & 1, grid, 2);
  }
  if (rc == LBM_OK) rc = batch_launch(s, p, 2, nsteps & 1, grid);
  return rc;
}

struct GraphHandle {
  hipGraphExec_t exec;
};

}  // namespace mlbm

using namespace mlbm;

extern "C" {

size_t lbm_batch_plan_size(int nitems) {
  if (nitems < 1) return 0;
  return kHeaderBytes + sizeof(BatchEntry) * (size_t)nitems;
}

int lbm_batch_plan_create(void* stream, const lbm_batch_item* items, int nitems,
                          int cells_per_thread, void* device_plan, lbm_batch_plan_t** plan,
                          long long* total_tiles) {
  g_last_error.clear();
  (void)hipGetLastError();
  if (!items || nitems < 1 || !device_plan || !plan)
    return fail(LBM_ERR_INVALID, "items, nitems >= 1, device_plan and plan needed");
  *plan = nullptr;
  const int cpt = cells_per_thread == 0 ? 1 : cells_per_thread;
  if (cpt < 1 || cpt > 16) return fail(LBM_ERR_INVALID, "cells_per_thread must be 1..16");
  const long long tile_cells = 256LL * cpt;
  // validate, then order the entries by (q, wall rule) so each class is a
  // contiguous run of entries and tiles
  std::vector<Geom> geo(nitems);
  for (int i = 0; i < nitems; ++i) {
    const lbm_batch_item& it = items[i];
    lbm_desc d = {};
    d.q = it.q;
    d.storage = LBM_F32;
    d.bb_mode = it.bb_mode;
    d.tau = it.tau;
    for (int k = 0; k < 3; ++k) d.dims[k] = it.dims[k];
    int rc = check_desc(&d, &geo[i]);
    if (rc) return fail(rc, "batch item " + std::to_string(i) + ": " + g_last_error);
    if (!it.f[0] || !it.f[1] || it.f[0] == it.f[1])
      return fail(LBM_ERR_INVALID, "batch item " + std::to_string(i) + ": needs two distinct state buffers");
    if (it.bb_mode != LBM_BB_NONE && !it.mask)
      return fail(LBM_ERR_INVALID, "batch item " + std::to_string(i) + ": bb_mode needs a mask");
    if (it.pitch != 0 && it.pitch < geo[i].plane)
      return fail(LBM_ERR_INVALID, "batch item " + std::to_string(i) + ": pitch is below the cell count");
    if (it.pitch > 0x7fffffffLL)
      return fail(LBM_ERR_INVALID, "batch item " + std::to_string(i) + ": pitch must fit int32");
  }
  // the wall-rule kernel of an item: halfway items with link words form
  // their own class (BB_HALF_LINKS)
  auto kernel_bb = [&](int i) {
    return (items[i].bb_mode == LBM_BB_HALFWAY && items[i].links) ? int(BB_HALF_LINKS)
                                                                   : items[i].bb_mode;
  };
  // periodic D3Q19 lattices the two-step march covers: rows of at most 341
  // cells, 19 planes of `pitch` cells addressable with 32-bit offsets
  auto pairable = [&](int i) {
    const long long pitch = items[i].pitch ? items[i].pitch : geo[i].plane;
    return items[i].q == 19 && items[i].bb_mode == LBM_BB_NONE && geo[i].nc <= 341 &&
                   19LL * pitch < (1LL << 31)
               ? 1
               : 0;
  };
  auto cls = [&](int i) { return (items[i].q * 4 + kernel_bb(i)) * 2 + pairable(i); };
  std::vector<int> order(nitems);
  for (int i = 0; i < nitems; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cls(x) < cls(y); });
  std::vector<char> host(lbm_batch_plan_size(nitems), 0);
  BatchEntry* out = reinterpret_cast<BatchEntry*>(host.data() + kHeaderBytes);
  lbm_batch_plan_t* hp = new lbm_batch_plan_t;
  hp->dplan = device_plan;
  long long tiles = 0, ptiles = 0;
  for (int j = 0; j < nitems; ++j) {
    const int i = order[j];
    const lbm_batch_item& it = items[i];
    const Geom& g = geo[i];
    BatchEntry& e = out[j];
    e.q = it.q;
    e.bb = kernel_bb(i);
    e.na = g.na; e.nb = g.nb; e.nc = g.nc;
    e.bpr = g.bpr;
    e.plane = g.plane;
    e.pitch = it.pitch ? it.pitch : g.plane;
    e.tile_begin = tiles;
    const int two = pairable(i);
    e.pair_begin = ptiles;
    e.tb = e.btiles = e.la = 0;
    long long np = 0;
    if (two) {
      e.tb = 1024 / g.nc - 2;
      e.btiles = (g.nb + e.tb - 1) / e.tb;
      e.la = kBatchPairChunk;
      np = (long long)e.btiles * ((g.na + e.la - 1) / e.la);
    }
    e.inv_tau = float(1.0 / it.tau);
    if (it.q == 19) {
      e.Fa = float(it.force[0]); e.Fb = float(it.force[1]); e.Fc = float(it.force[2]);
      e.lid_a = float(it.lid_velocity[0]); e.lid_b = float(it.lid_velocity[1]);
      e.lid_c = float(it.lid_velocity[2]);
    } else {
      e.Fa = 0.f; e.Fb = float(it.force[0]); e.Fc = float(it.force[1]);
      e.lid_a = 0.f; e.lid_b = float(it.lid_velocity[0]); e.lid_c = float(it.lid_velocity[1]);
    }
    e.buf[0] = it.f[0];
    e.buf[1] = it.f[1];
    e.mask = it.bb_mode == LBM_BB_NONE ? nullptr : it.mask;
    e.links = nullptr;
    if (it.bb_mode == LBM_BB_HALFWAY && it.links) {
      e.links = it.links;
      hp->links.push_back(LinkJob{it.q, it.mask, it.links, g});
    }
    e.rho = it.rho;
    e.u = it.u;
    e.p = it.p;
    const long long nt = (g.plane + tile_cells - 1) / tile_cells;
    if (hp->groups.empty() || hp->groups.back().q != it.q || hp->groups.back().bb != e.bb ||
        hp->groups.back().pair != two)
      hp->groups.push_back(BatchGroup{it.q, e.bb, two, j, j, tiles, tiles, ptiles, ptiles});
    hp->groups.back().e1 = j + 1;
    hp->groups.back().t1 = tiles + nt;
    hp->groups.back().p1 = ptiles + np;
    tiles += nt;
    ptiles += np;
  }
  for (const BatchGroup& gr : hp->groups) {
    if (gr.p1 - gr.p0 >= (1LL << 31)) {
      delete hp;
      return fail(LBM_ERR_UNSUPPORTED, "too many two-step workgroups in one lattice class");
    }
  }
  BatchHeader* hdr = reinterpret_cast<BatchHeader*>(host.data());
  hdr->total_tiles = tiles;
  hdr->nitems = nitems;
  hdr->cells_per_thread = cpt;
  if (total_tiles) *total_tiles = tiles;
  hipStream_t s = (hipStream_t)stream;
  hipError_t err = hipMemcpyAsync(device_plan, host.data(), host.size(), hipMemcpyHostToDevice, s);
  if (err == hipSuccess) err = hipStreamSynchronize(s);  // the host table dies with this call
  if (err != hipSuccess) {
    delete hp;
    return fail(LBM_ERR_HIP, std::string("plan upload: ") + hipGetErrorString(err));
  }
  *plan = hp;
  return LBM_OK;
}

int lbm_batch_plan_pairs(lbm_batch_plan_t* plan, int on) {
  g_last_error.clear();
  if (!plan) return fail(LBM_ERR_INVALID, "plan is NULL");
  plan->pairs = on ? 1 : 0;
  return LBM_OK;
}

int lbm_batch_plan_destroy(lbm_batch_plan_t* plan) {
  delete plan;
  return LBM_OK;
}

int lbm_batch_run(void* stream, const lbm_batch_plan_t* plan, int nsteps, int grid) {
  g_last_error.clear();
  (void)hipGetLastError();
  if (!plan || nsteps < 1 || grid < 1) return fail(LBM_ERR_INVALID, "plan, nsteps >= 1 and grid >= 1 needed");
  return batch_run((hipStream_t)stream, plan, nsteps, grid);
}

int lbm_batch_graph_create(const lbm_batch_plan_t* plan, int nsteps, int grid, void** graph) {
  g_last_error.clear();
  (void)hipGetLastError();
  if (!plan || nsteps < 1 || grid < 1 || !graph)
    return fail(LBM_ERR_INVALID, "plan, nsteps >= 1, grid >= 1 and graph needed");
  hipStream_t s;
  hipError_t err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (err != hipSuccess) return fail(LBM_ERR_HIP, std::string("stream: ") + hipGetErrorString(err));
  hipGraph_t g = nullptr;
  err = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  int rc = LBM_OK;
  if (err == hipSuccess) {
    rc = batch_run(s, plan, nsteps, grid);
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (err == hipSuccess) err = e2;
  }
  hipGraphExec_t exec = nullptr;
  if (err == hipSuccess && rc == LBM_OK) err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(s);
  if (rc != LBM_OK) return rc;
  if (err != hipSuccess) return fail(LBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(err));
  GraphHandle* hnd = new GraphHandle{exec};
  *graph = hnd;
  return LBM_OK;
}

int lbm_batch_graph_launch(void* stream, void* graph) {
  g_last_error.clear();
  if (!graph) return fail(LBM_ERR_INVALID, "graph is NULL");
  hipError_t err = hipGraphLaunch(static_cast<GraphHandle*>(graph)->exec, (hipStream_t)stream);
  if (err != hipSuccess) return fail(LBM_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(err));
  return LBM_OK;
}

int lbm_batch_graph_destroy(void* graph) {
  g_last_error.clear();
  if (!graph) return LBM_OK;
  GraphHandle* h = static_cast<GraphHandle*>(graph);
  hipError_t err = hipGraphExecDestroy(h->exec);
  delete h;
  if (err != hipSuccess) return fail(LBM_ERR_HIP, std::string("graph destroy: ") + hipGetErrorString(err));
  return LBM_OK;
}

}  // extern "C"
