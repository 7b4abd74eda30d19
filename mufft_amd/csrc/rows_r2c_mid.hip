// Instantiations of the fused row kernel: r2c (direction -1), the registered
// variants (row_variants.h) of 8192 and 16384 points.
#include "rows_impl.h"

namespace mufft {

void row_kernels_r2c_mid(const void **out) { fill_row_kernels<-1, KIND_R2C, kRowMidIndex, kRowSplitIndex>(out); }

}  // namespace mufft
