// Instantiations of the fused row kernel: c2c_inv (direction 1), every
// registered variant (row_variants.h).
#include "rows_impl.h"

namespace mufft {

void row_kernels_c2c_inv(const void **out) { fill_row_kernels<1, KIND_C2C>(out); }

}  // namespace mufft
