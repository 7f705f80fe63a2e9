// Kernel instantiations of the fused step for D3Q19.
#include "lbm_step_impl.h"

namespace lbm {
KernelFn pick_kernel_d3q19(int storage, int coll, bool mask, int mode) {
  return pick_kernel<D3Q19>(storage, coll, mask, mode);
}
}  // namespace lbm
