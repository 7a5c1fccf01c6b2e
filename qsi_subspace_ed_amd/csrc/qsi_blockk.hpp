// Block apply shared by qsi_op_apply_block, the FTLM block Lanczos and the
// solve engine: many columns (start vectors or requests) over one or more
// operators, rows cut into 64-row tiles walked by a persistent grid of
// wavefronts.  The kernels live in qsi_block.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "qsi_device.hpp"
#include "qsi_internal.hpp"
#include "qsi_reduce.hpp"

namespace qsi {

constexpr int kBlkThreads = 256;
constexpr int kBlkWaves = kBlkThreads / 64;
constexpr int kTileRows = 64;
constexpr int kMaxSectors = 8;

struct BlkSector {
  BasisView b;
  const OffTerm* off;
  const DiagTerm* dg;
  const double* coeffs;  // n_slots x R, or null
  const CoefClass* cls;  // coefficient classes of the off-diagonal terms
  int n_cls;             // 0: per-term coefficients
  int n_off;
  int n_diag;
  int64_t tile_begin;  // first global tile of this sector
  int64_t row_off;     // first row of this sector in the concatenated blocks
  // cached connectivity of the operator, or null (see qsi_op_s)
  const uint64_t* rowptr;
  const uint32_t* cj;
  const uint8_t* cm;
};

struct BlkArgs {
  BlkSector sec[kMaxSectors];
  int n_sec;
  int64_t n_tiles;
  int R;
  int G;            // lanes per row group: a power of two >= R, or any width
                    // >= R for the cached kernel without partials (the
                    // column-pairs kernel sets its own, ceil(R / 2))
  int64_t ldx, ldy; // row strides of the X and Y blocks (>= R)
  int accumulate;   // Y += H X instead of Y = H X (apply kernel)
  int y_colmajor;   // Y stored column-major: Y[col * ldy + row] (apply kernel)
  int cfg;          // kernel configuration (blk_cfg_*), chosen by the autotuner
};

// Block-apply configuration: bits 0-1 pick the kernel (cached connectivity:
// 4 / 8 / 16 gathers in flight per lane; searching: 4 / 4 / 2 lookups in
// flight), bit 2 halves the persistent grid, bit 3 lets the cached kernel
// serve two columns per lane (16-byte gathers) where X's rows allow it.
constexpr int kBlkCfgDefault = 1;
inline int blk_cfg_kernel(int cfg) { return cfg & 3; }
inline bool blk_cfg_half(int cfg) { return (cfg >> 2) & 1; }
inline bool blk_cfg_pairs(int cfg) { return (cfg >> 3) & 1; }

// True when every sector of `a` has its connectivity cached: the block
// apply then gathers through the cached ranks instead of searching.
inline bool blk_cached(const BlkArgs& a) {
  for (int k = 0; k < a.n_sec; ++k)
    if (!a.sec[k].rowptr) return false;
  return a.n_sec > 0;
}

// True when a cached block apply on X serves two columns per lane: the
// configuration's pairs bit, R >= 2, an even ldx >= R rounded up to even and
// a 16-byte aligned X.
bool blk_pairs(const BlkArgs& a, const double* X);

// Blocks of a persistent block-apply grid over n_tiles 64-row tiles: all
// resident at once (occupancy x compute units), never more than the tiles
// need.  `cached` picks the kernel (see blk_cached).
int blk_apply_blocks(int nw, int64_t n_tiles, bool cached, int cfg);

// Y (+)= H X over the sectors of `a` (the block-apply kernels; see
// qsi_block.hip).  partials: n_sec x (n_blocks * kBlkWaves) x R, or null.
int launch_blk_apply(const BlkArgs& a, int nw, int n_blocks, const double* X, double* Y,
                     double* partials, hipStream_t s);

// Builds (enable != 0) or frees the operator's cached connectivity.
int op_cache_connectivity(qsi_op_s* op, bool enable, bool allow_pad = true);

}  // namespace qsi
