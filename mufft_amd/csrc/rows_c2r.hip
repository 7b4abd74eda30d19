// Instantiations of the fused row kernel: c2r (direction 1), the registered
// variants (row_variants.h) below 8192 points.
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2r(const void **out) { fill_row_kernels<1, KIND_C2R, 0, kRowMidIndex>(out); }

}  // namespace mufft
