// Instantiations of the fused row kernel: c2r (direction 1), the registered
// variants (row_variants.h) of 8192 and 16384 points.
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2r_mid(const void **out) { fill_row_kernels<1, KIND_C2R, kRowMidIndex, kRowSplitIndex>(out); }

}  // namespace mufft
