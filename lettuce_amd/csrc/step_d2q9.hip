// Kernel instantiations of the fused step for D2Q9.
#include "lbm_step_impl.h"

namespace lbm {
KernelFn pick_kernel_d2q9(int storage, int coll, bool mask, int mode) {
  return pick_kernel<D2Q9>(storage, coll, mask, mode);
}
}  // namespace lbm
