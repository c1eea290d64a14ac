// Host-side helpers shared by the C ABI translation units: error state,
// descriptor validation and launch geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/lbm_c_api.h"

namespace mlbm {

// BB_HALF_LINKS: halfway bounce-back whose pull gathers read per-cell link
// words (FwdArgs::links) instead of the neighbours' mask bytes; a separate
// instantiation, so neither gather carries the other's registers.
enum { BB_NONE = 0, BB_REF = 1, BB_HALF = 2, BB_HALF_LINKS = 3 };

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

// XCD runs (xcd_remap) for row-block launches. Measured on MI355X, D3Q19
// f32, per fused launch in TB/s, identity -> runs of 512 row-blocks per XCD:
// lattices whose a-plane (nb * nc cells) is a power of two gain while the
// grid is small: 64x128x256 4.38 -> 5.30, 128x128x256 4.89 -> 5.62,
// 160x128x256 5.23 -> 5.57, 128^3 4.53 -> 5.05, D2Q9 2048^2 4.51 -> 5.48;
// above 24576 row-blocks they lose (224x128x256 5.65 -> 5.51, 256^3 5.71 ->
// 5.16). Other shapes lose at any size: 100x100x256 5.31 -> 5.19,
// 128x128x192 5.22 -> 4.93, 64x300x256 5.48 -> 5.27. Shorter runs (2-64)
// stay within a few per cent of identity everywhere.
constexpr int kAutoXcdMaxBlocks = 24576;
constexpr int kAutoXcdGroup = 512;

inline bool pow2(long long v) { return v > 0 && (v & (v - 1)) == 0; }

// Run length G of the XCD remap for a row-block launch over g (0: identity).
// Without LBM_FLAG_XCD_REMAP the measured policy above applies; with it,
// d->xcd_group (1 is the identity, 0 one contiguous eighth of the lattice
// per XCD).
inline int xcd_group(const lbm_desc* d, const Geom& g) {
  if (!(d->flags & LBM_FLAG_XCD_REMAP))
    return (g.nblocks <= kAutoXcdMaxBlocks && pow2((long long)g.nb * g.nc)) ? kAutoXcdGroup : 0;
  return d->xcd_group > 0 ? d->xcd_group : g.nblocks / 8;
}

// Row blocks leave rows that are not a multiple of 64 cells misaligned, and
// rows that are not a multiple of 256 cells end in a partly idle workgroup
// (or use workgroups of 1-3 waves). Measured on MI355X, D3Q19 f32 per fused
// launch, row blocks -> 256-cell tiles of the flattened grid, TB/s:
// 256x160x161 3.06 -> 5.41, 256x160x193 3.36 -> 5.50, 321x256x255 3.71 ->
// 5.64, 100^3 3.78 -> 4.99, 300x300x64 5.25 -> 5.65, 256x256x128 5.37 ->
// 5.56, 150x150x64 4.89 -> 5.09; rows of 192 come out even (5.27 -> 5.11,
// 4.89 -> 4.86, 5.23 -> 5.25) and rows of 256 too (5.64 -> 5.63). A plane
// size that is not a multiple of 64 keeps most stores misaligned either way
// (161^3 3.12 -> 3.48). So row blocks are kept for rows that are a multiple
// of 256 cells, and everything else runs over flattened tiles, in identity
// order: runs of 512 there went from +9 % (128x128x192) to -23 % (160^3).
// block_x > 0 asks for row blocks of that width, -1 for flattened tiles.
inline int flat_tiles(const lbm_desc* d, const Geom& g) {
  if (d->block_x > 0) return 0;
  if (d->block_x == 0 && g.nc % 256 == 0) return 0;
  return (int)((g.plane + 255) / 256);
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
  if (bx == 0 || bx == -1) {
    bx = g->nc >= 256 ? 256 : ((g->nc + 63) / 64) * 64;
  }
  if (bx < 64 || bx > 256 || bx % 64 != 0)
    return fail(LBM_ERR_INVALID, "block_x must be 0, -1, 64, 128, 192 or 256");
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
