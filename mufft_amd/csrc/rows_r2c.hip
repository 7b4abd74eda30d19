// Instantiations of the fused row kernel: r2c (direction -1), the registered
// variants (row_variants.h) below 8192 points.
#include "rows_impl.h"

namespace mufft {

void row_kernels_r2c(const void **out) { fill_row_kernels<-1, KIND_R2C, 0, kRowMidIndex>(out); }

}  // namespace mufft
