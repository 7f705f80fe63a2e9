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

}  // namespace lbm
