// Instantiations of the fused row kernel: r2c (direction -1), every
// registered variant (row_variants.h).
#include "rows_impl.h"

namespace mufft {

void row_kernels_r2c(const void **out) { fill_row_kernels<-1, KIND_R2C>(out); }

}  // namespace mufft
