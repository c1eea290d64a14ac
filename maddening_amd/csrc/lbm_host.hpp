// Host-side helpers shared by the C ABI translation units: error state,
// descriptor validation and launch geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lbm_c_api.h"

namespace mlbm {

enum { BB_NONE = 0, BB_REF = 1, BB_HALF = 2 };

// message for lbm_last_error(), one per host thread
inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

struct Geom {
  int na, nb, nc;
  long long plane;
  int bx, bpr, nblocks;
  int rowv = 0;  // vector width of the row-per-wave kernel, 0 = per-cell kernel
};

// Lattices of up to this many row-blocks run faster with runs of
// kAutoXcdGroup consecutive row-blocks per XCD. Measured on MI355X, D3Q19
// f32 with 256-thread blocks, per fused launch in TB/s, identity -> runs of 512:
// 8192 blocks 4.38 -> 5.30, 16384 4.89 -> 5.62, 20480 5.23 -> 5.57,
// 24576 5.49 -> 5.54, 28672 5.65 -> 5.51, 65536 (256^3) 5.71 -> 5.16;
// D2Q9 2048^2 (16384 blocks) 4.51 -> 5.48.
constexpr int kAutoXcdMaxBlocks = 24576;
constexpr int kAutoXcdGroup = 512;

// Run length G of the XCD remap for a launch over g.nblocks row-blocks
// (0: identity). Without LBM_FLAG_XCD_REMAP the measured policy above
// applies; with it, d->xcd_group (1 is the identity, 0 one contiguous eighth
// of the lattice per XCD).
inline int xcd_group(const lbm_desc* d, const Geom& g) {
  if (!(d->flags & LBM_FLAG_XCD_REMAP))
    return g.nblocks <= kAutoXcdMaxBlocks ? kAutoXcdGroup : 0;
  return d->xcd_group > 0 ? d->xcd_group : g.nblocks / 8;
}

inline int storage_size(int s) {
  switch (s) {
    case LBM_F32: return 4;
    case LBM_F64: return 8;
    case LBM_F16_DEV: return 2;
    case LBM_BF16_DEV: return 2;
    case LBM_F32_DEV: return 4;
    default: return 0;
  }
}

inline int check_desc(const lbm_desc* d, Geom* g) {
  if (!d) return fail(LBM_ERR_INVALID, "lbm_desc is NULL");
  if (d->q != 9 && d->q != 19)
    return fail(LBM_ERR_INVALID, "q must be 9 (D2Q9) or 19 (D3Q19), got " + std::to_string(d->q));
  if (storage_size(d->storage) == 0)
    return fail(LBM_ERR_INVALID, "unknown storage type " + std::to_string(d->storage));
  if (d->bb_mode < 0 || d->bb_mode > 2)
    return fail(LBM_ERR_INVALID, "unknown bb_mode " + std::to_string(d->bb_mode));
  if (d->scheme < 0 || d->scheme > 2)
    return fail(LBM_ERR_INVALID, "unknown scheme " + std::to_string(d->scheme));
  if (!(d->tau > 0.5))
    return fail(LBM_ERR_INVALID, "tau must be > 0.5 (got " + std::to_string(d->tau) + ")");
  long long n0 = d->dims[0], n1 = d->dims[1], n2 = d->q == 19 ? d->dims[2] : 1;
  if (n0 < 1 || n1 < 1 || n2 < 1) return fail(LBM_ERR_INVALID, "grid dims must be >= 1");
  if (d->q == 19) {
    g->na = (int)n0; g->nb = (int)n1; g->nc = (int)n2;
  } else {
    g->na = 1; g->nb = (int)n0; g->nc = (int)n1;
  }
  const long long plane = (long long)g->na * g->nb * g->nc;
  if (plane >= (1LL << 31) || n0 >= (1LL << 31) || n1 >= (1LL << 31) || n2 >= (1LL << 31))
    return fail(LBM_ERR_UNSUPPORTED, "more than 2^31-1 cells per population plane");
  g->plane = plane;
  int bx = d->block_x;
  if (bx == 0) {
    bx = g->nc >= 256 ? 256 : ((g->nc + 63) / 64) * 64;
  }
  if (bx < 64 || bx > 256 || bx % 64 != 0)
    return fail(LBM_ERR_INVALID, "block_x must be 64, 128, 192 or 256");
  if (d->row_vec != 0 && d->row_vec != -1 && d->row_vec != 1 && d->row_vec != 2 && d->row_vec != 4)
    return fail(LBM_ERR_INVALID, "row_vec must be -1, 0, 1, 2 or 4");
  if (d->row_vec == 4 && d->storage == LBM_F64)
    return fail(LBM_ERR_UNSUPPORTED, "row_vec 4 is not built for f64 storage (use 1 or 2)");
  if (d->row_vec > 0 && d->storage == LBM_F32_DEV)
    return fail(LBM_ERR_UNSUPPORTED, "the row-per-wave kernels are not built for f32dev storage");
  const int D = d->q == 19 ? 3 : 2;
  for (int face : {d->inlet_face, d->outlet_face}) {
    if (face < 0 || face > 6)
      return fail(LBM_ERR_INVALID, "inlet_face/outlet_face must be 0 (none) or 1..6");
    if (face > 0 && (face - 1) / 2 >= D)
      return fail(LBM_ERR_INVALID, "a pressure face uses an axis the lattice does not have");
  }
  if (d->row_vec > 0 && g->nc % (64 * d->row_vec) != 0)
    return fail(LBM_ERR_INVALID, "row_vec needs the contiguous axis to be a multiple of 64*row_vec");
  g->bx = bx;
  g->bpr = (g->nc + bx - 1) / bx;
  const long long nblocks = (long long)g->na * g->nb * g->bpr;
  if (nblocks >= (1LL << 31)) return fail(LBM_ERR_UNSUPPORTED, "grid too large");
  g->nblocks = (int)nblocks;
  return LBM_OK;
}

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return LBM_OK;
}

}  // namespace mlbm
