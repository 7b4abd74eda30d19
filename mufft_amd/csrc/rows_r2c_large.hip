// Instantiations of the fused row kernel: r2c (direction -1), the registered
// variants (row_variants.h) from 32768 points on.
#include "rows_impl.h"

namespace mufft {

void row_kernels_r2c_large(const void **out) { fill_row_kernels<-1, KIND_R2C, kRowSplitIndex, kNumRowVariants>(out); }

}  // namespace mufft
