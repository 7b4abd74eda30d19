// Instantiations of the fused row kernel: c2c forward (direction -1), the registered
// variants (row_variants.h) below 8192 points.
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2c_fwd(const void **out) { fill_row_kernels<-1, KIND_C2C, 0, kRowMidIndex>(out); }

}  // namespace mufft
