// Forward-direction instantiations of the fused row kernel (c2c and r2c).
#include "rows_impl.h"

namespace mufft {

const void *row_kernel_fwd(int logN, int kind) {
    if (kind == KIND_C2C) return row_kernel_ptr<-1, KIND_C2C>(logN);
    if (kind == KIND_R2C) return row_kernel_ptr<-1, KIND_R2C>(logN);
    return nullptr;
}

}  // namespace mufft
