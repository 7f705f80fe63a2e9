// Shared declarations of the two-step (temporally blocked) kernel.
#pragma once

#include "lbm_step_impl.h"

namespace lbm {

constexpr int TB_Y = 8;                        // interior rows per block
constexpr int TB_ROWS = TB_Y + 2;              // with one halo row per side
constexpr int TB_THREADS = 64 * (TB_ROWS + 1); // row wavefronts + z-edge wavefront

struct TbParams {
  KParams k;
  int cx;       // planes per block along x
  int nzt;      // tiles along z (n2 / 64)
  int nyt;      // tiles along y (n1 / TB_Y)
  int ntiles;   // nzt * nyt * x chunks
  int per_xcd;  // 0: block b takes tile b; else tiles per XCD (see tb2_kernel)
};
using TbFn = void (*)(const TbParams);

// two-step kernel for (stencil, storage, collision), or nullptr, with one
// or two block-wide barriers per plane; wide: for arrays whose offsets do
// not fit 32 bits (one barrier per plane)
TbFn pick_tb2_kernel(int stencil, int storage, int coll, int barriers, bool wide);

// The two-step kernel for 2D grids, with or without a cell-class map
// (lbm_tb2d.hip):
// one block of nr x n2 threads (nr = 1 or 2 rows per iteration) per chunk
// of cx rows, with tb2d_lds_bytes of dynamic LDS (the S1 ring) besides
// tb2d_static_lds_bytes of its own.
struct Tb2dParams {
  KParams k;
  int cx;  // rows (i1) per block
};
using Tb2dFn = void (*)(const Tb2dParams);
// kernel for (stencil, storage codec, collision, nr, with a class map), or
// nullptr
Tb2dFn pick_tb2d_kernel(int stencil, int codec, int coll, int nr, bool mask);
// dynamic LDS bytes of a block for rows of n2 cells, storage_bytes per
// population and nr rows per iteration, and the kernel's static LDS bytes
// (class fields)
int64_t tb2d_lds_bytes(int n2, int storage_bytes, int nr);
int64_t tb2d_static_lds_bytes();
// rows of S1 beyond each end of a block's chunk
int tb2d_halo();

}  // namespace lbm
