// Instantiations of the fused row kernel: c2r (direction 1), every
// registered variant (row_variants.h).
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2r(const void **out) { fill_row_kernels<1, KIND_C2R>(out); }

}  // namespace mufft
