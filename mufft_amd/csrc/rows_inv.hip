// Inverse-direction instantiations of the fused row kernel (c2c and c2r).
#include "rows_impl.h"

namespace mufft {

const void *row_kernel_inv(int logN, int kind) {
    if (kind == KIND_C2C) return row_kernel_ptr<1, KIND_C2C>(logN);
    if (kind == KIND_C2R) return row_kernel_ptr<1, KIND_C2R>(logN);
    return nullptr;
}

}  // namespace mufft
