// Instantiations of the fused row kernel: c2c forward (direction -1), the registered
// variants (row_variants.h) from 32768 points on.
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2c_fwd_large(const void **out) { fill_row_kernels<-1, KIND_C2C, kRowSplitIndex, kNumRowVariants>(out); }

}  // namespace mufft
