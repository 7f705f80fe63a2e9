// Kernel instantiations of the fused step for D3Q27.
#include "lbm_step_impl.h"

namespace lbm {
KernelFn pick_kernel_d3q27(int storage, int coll, bool mask, int mode) {
  return pick_kernel<D3Q27>(storage, coll, mask, mode);
}
KernelFn pick_pair_kernel_d3q27(int storage, int coll, int mode, bool nt) {
  return pick_pair_kernel<D3Q27>(storage, coll, mode, nt);
}
KernelFn pick_row_kernel_d3q27(int storage, int coll) {
  return pick_row_kernel<D3Q27>(storage, coll);
}
}  // namespace lbm
