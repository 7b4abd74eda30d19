// Instantiations of the fused row kernel: c2r (direction 1), the registered
// variants (row_variants.h) from 32768 points on.
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2r_large(const void **out) { fill_row_kernels<1, KIND_C2R, kRowSplitIndex, kNumRowVariants>(out); }

}  // namespace mufft
