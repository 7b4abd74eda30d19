// Convolution plans (fft.c:549-624, 794-809): two input transforms, the
// multiply fused into the inverse transform, and the batched halves.
#include <new>

#include "host.h"

using namespace mufft;

extern "C" {

void mufft_free_plan_conv(mufft_plan_conv *plan) {
    if (!plan) return;
    destroy_1d(plan->plans[0]);
    destroy_1d(plan->plans[1]);
    destroy_1d(plan->output_plan);
    delete plan;
}

static mufft_plan_conv *mufft_create_plan_conv_impl(unsigned N, unsigned flags, unsigned method) {
    if (!is_pow2(N) || N < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    mufft_plan_conv *c = new (std::nothrow) mufft_plan_conv;
    if (!c) return nullptr;
    const unsigned ef = (method & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST) ? MUFFT_FLAG_ZERO_PAD_UPPER_HALF : 0u;
    const unsigned es = (method & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND) ? MUFFT_FLAG_ZERO_PAD_UPPER_HALF : 0u;
    c->N = N;
    c->method = method;
    c->stereo = (method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
    if (!c->stereo) {
        c->block_size = (N / 2 + 4) * 8;   // fft.c:586 (MUFFT_PADDING_COMPLEX_SAMPLES = 4)
        c->plans[0] = mufft_create_plan_1d_r2c(N, flags | ef);
        c->plans[1] = mufft_create_plan_1d_r2c(N, flags | es);
        c->output_plan = mufft_create_plan_1d_c2r(N, flags);
        c->conv_multiply_n = N / 2 + 1;
    } else {
        c->block_size = (size_t)N * 8;
        c->plans[0] = mufft_create_plan_1d_c2c(N, MUFFT_FORWARD, flags | ef);
        c->plans[1] = mufft_create_plan_1d_r2c(N, flags | es | MUFFT_FLAG_FULL_R2C);
        c->output_plan = mufft_create_plan_1d_c2c(N, MUFFT_INVERSE, flags);
        c->conv_multiply_n = N;
    }
    c->normalization = 1.0f / (float)N;
    if (!c->plans[0] || !c->plans[1] || !c->output_plan) {
        mufft_free_plan_conv(c);
        return nullptr;
    }
    return c;
}


mufft_plan_conv *mufft_create_plan_conv(unsigned N, unsigned flags, unsigned method) {
    return guarded<mufft_plan_conv *>(nullptr, [&] { return mufft_create_plan_conv_impl(N, flags, method); });
}

size_t mufft_conv_get_transformed_block_size(mufft_plan_conv *plan) {
    return plan ? plan->block_size : 0;
}

void mufft_execute_conv_input(mufft_plan_conv *plan, unsigned block, void *output, const void *input) {
    if (!plan || block > 1) {
        set_err("invalid convolution block");
        return;
    }
    mufft_execute_plan_1d(plan->plans[block], output, input);
}

void mufft_execute_conv_output(mufft_plan_conv *plan, void *output, const void *input_first,
                               const void *input_second) {
    if (!plan) return;
    // The reference multiplies into conv_block and then runs the inverse plan
    // (fft.c:804-809); here the multiply is the inverse kernel's prologue.
    mufft_plan_1d *op = plan->output_plan;
    const size_t nb = (size_t)plan->conv_multiply_n * 8;
    exec_1d_dropin(op, output, input_first, input_second, nb, RF_MUL, plan->normalization);
}

int mufft_hip_execute_conv_input_batched(mufft_plan_conv *plan, unsigned block, void *output,
                                         const void *input, unsigned howmany, size_t idist,
                                         size_t odist, void *stream) {
    if (!plan || block > 1) {
        set_err("invalid convolution block");
        return -1;
    }
    mufft_plan_1d *p = plan->plans[block];
    const size_t od = odist ? odist : plan->block_size / 8;
    // Input distances of real inputs are in floats (mufft_hip_execute_plan_1d_batched).
    return mufft_hip_execute_plan_1d_batched(p, output, input, howmany, idist, od, stream);
}

int mufft_hip_execute_conv_output_batched(mufft_plan_conv *plan, void *output,
                                          const void *first, const void *second,
                                          unsigned howmany, size_t odist, size_t dist_first,
                                          size_t dist_second, void *stream) {
    if (!plan || !output || !first || !second) {
        set_err("null argument");
        return -1;
    }
    mufft_plan_1d *op = plan->output_plan;
    const long long bd = (long long)(plan->block_size / 8);
    const long long d1 = dist_first ? (long long)dist_first : bd;
    const long long d2 = dist_second ? (long long)dist_second : bd;
    long long od;
    if (op->kind == KIND_C2R) {
        if (odist & 1) {
            set_err("real output distance must be even");
            return -1;
        }
        od = odist ? (long long)odist / 2 : (long long)op->N;
    } else {
        od = odist ? (long long)odist : (long long)op->N;
    }
    if (kernel_ptrs(&output, &first, &second)) return -1;
    return exec_1d(op, output, first, second, howmany, d1, d2, od, RF_MUL, plan->normalization,
                   static_cast<hipStream_t>(stream));
}

}  // extern "C"

