// 1D plans: creation (fft.c:476-662), execution (fft.c:811-850) and the
// batched, adjoint and fused-multiply extensions.
#include <stdio.h>

#include <new>

#include "host.h"
#include "staging.h"

namespace mufft {

void destroy_1d(mufft_plan_1d *p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    p->eng.release();
    if (p->rtw) (void)hipFree(p->rtw);
    p->adj.release();
    if (p->adj_rtw) (void)hipFree(p->adj_rtw);
    p->fold.release();
    p->s0.release();
    p->s1.release();
    p->stage_in.release();
    p->stage_in2.release();
    p->stage_out.release();
    free_pipe(p->pipe);
    p->scratch.release();
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static mufft_plan_1d *create_1d_impl(int kind, unsigned Nc, int direction, unsigned flags, int var) {
    if (!is_pow2(Nc) || Nc < 2 || Nc > (1u << 27)) {
        set_err("invalid transform size");
        return nullptr;
    }
    const int logn = ilog2(Nc);
    if (var != 0 && (logn > ROW_MAX_LOG2 || !row_variant_registered(logn, var))) {
        set_err("variant %#x is not registered for %u points", var, Nc);
        return nullptr;
    }
    mufft_plan_1d *p = new (std::nothrow) mufft_plan_1d;
    if (!p) return nullptr;
    p->N = Nc;
    p->kind = kind;
    p->direction = direction;
    p->flags = flags;
    p->full = (kind == KIND_R2C) && (flags & MUFFT_FLAG_FULL_R2C);
    p->zero_pad = (flags & MUFFT_FLAG_ZERO_PAD_UPPER_HALF) != 0;
    if (hipGetDevice(&p->device) != hipSuccess ||
        hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_err("no usable HIP device");
        p->stream = nullptr;
        destroy_1d(p);
        return nullptr;
    }
    if (p->eng.init((int)Nc, direction, kind, true, false, p->stream, var)) {
        destroy_1d(p);
        return nullptr;
    }
    if (kind != KIND_C2C && build_resolve_table((int)Nc, direction, &p->rtw, p->stream)) {
        destroy_1d(p);
        return nullptr;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        set_err("twiddle generation failed: %s", hipGetErrorString(hipGetLastError()));
        destroy_1d(p);
        return nullptr;
    }
    const size_t n = Nc;
    switch (kind) {
    case KIND_C2C:
        p->in_bytes = (p->zero_pad ? n / 2 : n) * 8;
        p->out_bytes = n * 8;
        break;
    case KIND_R2C:
        p->in_bytes = (p->zero_pad ? n : 2 * n) * 4;
        p->out_bytes = (p->full ? 2 * n : n + 1) * 8;
        break;
    default:
        p->in_bytes = (n + 1) * 8;
        p->out_bytes = 2 * n * 4;
        break;
    }
    return p;
}

mufft_plan_1d *create_1d(int kind, unsigned Nc, int direction, unsigned flags, int var) {
    return guarded<mufft_plan_1d *>(nullptr, [&] { return create_1d_impl(kind, Nc, direction, flags, var); });
}

int exec_1d(mufft_plan_1d *p, void *out, const void *in, const void *in2, unsigned howmany,
            long long idist, long long idist2, long long odist, unsigned extra, float norm,
            hipStream_t st) {
    unsigned flags = extra;
    if (p->zero_pad) flags |= RF_ZERO_PAD;
    if (p->full) flags |= RF_FULL;
    auto run = [&] {
        return run_rows(p->eng, p->kind, p->rtw, static_cast<const float2 *>(in), idist,
                        static_cast<const float2 *>(in2), idist2, static_cast<float2 *>(out), odist,
                        (int)howmany, flags, norm, p->s0, p->s1, st);
    };
    // The fused kernels touch no plan-owned buffer but the read-only tables.
    if (!rows_use_scratch(p->eng)) return run();
    return with_scratch(p->scratch, st, run);
}

long long in_dist_c(const mufft_plan_1d *p) {
    if (p->kind == KIND_C2R) return (long long)p->N + 1;
    return p->zero_pad ? (long long)p->N / 2 : (long long)p->N;
}

long long out_dist_c(const mufft_plan_1d *p) {
    if (p->kind == KIND_R2C) return p->full ? 2LL * p->N : (long long)p->N + 1;
    return p->N;
}

// Kind / direction / weighting of the adjoint of a plan's transform:
// c2c -> opposite-direction c2c; r2c -> c2r with RF_ADJ_R2C; c2r -> r2c with RF_ADJ_C2R.
static void adjoint_shape(const mufft_plan_1d *p, int *kind, int *dir, unsigned *flags) {
    if (p->kind == KIND_C2C) {
        *kind = KIND_C2C;
        *dir = -p->direction;
        *flags = 0;
    } else if (p->kind == KIND_R2C) {
        *kind = KIND_C2R;
        *dir = MUFFT_INVERSE;
        *flags = RF_ADJ_R2C;
    } else {
        *kind = KIND_R2C;
        *dir = MUFFT_FORWARD;
        *flags = RF_ADJ_C2R;
    }
}

// Builds the adjoint engine on first use; callers hold p->scratch.m.
static int ensure_adjoint(mufft_plan_1d *p) {
    if (p->adj_ready) return 0;
    if (p->kind == KIND_C2R && p->zero_pad) {
        set_err("no adjoint for zero-padded c2r plans");
        return -1;
    }
    int kind, dir;
    unsigned fl;
    adjoint_shape(p, &kind, &dir, &fl);
    if (p->adj.init((int)p->N, dir, kind, true, false, p->stream)) return -1;
    if (kind != KIND_C2C && build_resolve_table((int)p->N, dir, &p->adj_rtw, p->stream)) return -1;
    HIP_TRY(hipStreamSynchronize(p->stream));
    p->adj_ready = true;
    return 0;
}

// Adjoint execution; distances in complex elements of the adjoint's input
// (the forward output) and output (the forward input).
static int exec_adjoint(mufft_plan_1d *p, void *out, const void *in, const void *in2,
                        unsigned howmany, long long idist, long long idist2, long long odist,
                        unsigned extra, float norm, hipStream_t st) {
    return with_scratch(p->scratch, st, [&]() -> int {
        if (ensure_adjoint(p)) return -1;
        int kind, dir;
        unsigned fl;
        adjoint_shape(p, &kind, &dir, &fl);
        const float2 *src = static_cast<const float2 *>(in);
        long long sd = idist;
        if (p->kind == KIND_R2C && p->full) {
            if (extra & RF_MUL) {
                set_err("fused multiply is not available for the adjoint of a full r2c plan");
                return -1;
            }
            // Full spectrum (2N bins) folded onto bins 0..N first.
            const long long nb = (long long)p->N + 1;
            if (p->fold.ensure(sizeof(float2) * (size_t)nb * howmany)) return -1;
            MulArgs m{};
            m.a = src;
            m.out = p->fold.f2();
            m.n = 2LL * p->N;
            m.adist = idist;
            m.odist = nb;
            m.batch = (int)howmany;
            HIP_TRY(launch_fold(m, st));
            src = p->fold.f2();
            sd = nb;
        }
        const float2 *rtw = kind == KIND_C2C ? nullptr : p->adj_rtw;
        return run_rows(p->adj, kind, rtw, src, sd, static_cast<const float2 *>(in2), idist2,
                        static_cast<float2 *>(out), odist, (int)howmany, fl | extra, norm, p->s0,
                        p->s1, st);
    });
}

void exec_1d_dropin(mufft_plan_1d *p, void *out, const void *in, const void *in2, size_t in2_bytes,
                    unsigned extra, float norm) {
    std::lock_guard<std::mutex> g(p->lock);
    int dev_prev = 0;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != p->device) (void)hipSetDevice(p->device);
    const void *din = in, *din2 = in2;
    void *dout = out;
    const PtrInfo ci = classify_ptr(in), co = classify_ptr(out);
    const PtrInfo ci2 = in2 ? classify_ptr(in2) : PtrInfo{PTR_DEVICE, const_cast<void *>(in2)};
    // One fused launch touches each point once: pinned host buffers are then
    // read and written in place over PCIe, with no staging copies.
    const bool direct = p->eng.row_fused() && ci.kind != PTR_PAGEABLE && co.kind != PTR_PAGEABLE &&
                        ci2.kind != PTR_PAGEABLE;
    const bool hin = !direct && ci.kind != PTR_DEVICE, hout = !direct && co.kind != PTR_DEVICE;
    const bool hin2 = !direct && in2 && ci2.kind != PTR_DEVICE;
    if (direct) {
        din = ci.dev;
        din2 = ci2.dev;
        dout = co.dev;
    }
    int rc = 0;
    if (hin) {
        rc |= p->stage_in.ensure(p->in_bytes);
        if (!rc) rc |= hipMemcpyAsync(p->stage_in.p, in, p->in_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
        din = p->stage_in.p;
    }
    if (!rc && hin2) {
        rc |= p->stage_in2.ensure(in2_bytes);
        if (!rc) rc |= hipMemcpyAsync(p->stage_in2.p, in2, in2_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
        din2 = p->stage_in2.p;
    }
    if (!rc && hout) {
        rc |= p->stage_out.ensure(p->out_bytes);
        dout = p->stage_out.p;
    }
    if (!rc) rc = exec_1d(p, dout, din, din2, 1, in_dist_c(p), in_dist_c(p), out_dist_c(p), extra, norm, p->stream);
    if (!rc && hout)
        rc |= hipMemcpyAsync(out, p->stage_out.p, p->out_bytes, hipMemcpyDeviceToHost, p->stream) != hipSuccess;
    if (hipStreamSynchronize(p->stream) != hipSuccess || rc) {
        if (!rc) set_err("execution failed: %s", hipGetErrorString(hipGetLastError()));
        fprintf(stderr, "mufft_hip: %s\n", g_err);
    }
    if (dev_prev != p->device) (void)hipSetDevice(dev_prev);
}

// Converts API distances (elements of each buffer's type) to complex elements.
static int api_dists(const mufft_plan_1d *p, bool adjoint, size_t idist, size_t odist,
                     long long *id, long long *od) {
    // Buffer types: forward c2c C->C, r2c R->C, c2r C->R; the adjoint swaps them.
    const bool in_real = adjoint ? (p->kind == KIND_C2R) : (p->kind == KIND_R2C);
    const bool out_real = adjoint ? (p->kind == KIND_R2C) : (p->kind == KIND_C2R);
    const long long in_packed = adjoint ? (p->kind == KIND_C2R ? (long long)p->N : out_dist_c(p))
                                        : in_dist_c(p);
    const long long out_packed = adjoint ? (p->kind == KIND_R2C ? (long long)p->N
                                            : (p->kind == KIND_C2R ? (long long)p->N + 1 : (long long)p->N))
                                         : out_dist_c(p);
    if ((in_real && (idist & 1)) || (out_real && (odist & 1))) {
        set_err("real buffer distances must be even");
        return -1;
    }
    *id = idist ? (long long)(in_real ? idist / 2 : idist) : in_packed;
    *od = odist ? (long long)(out_real ? odist / 2 : odist) : out_packed;
    return 0;
}

}  // namespace mufft

using namespace mufft;

extern "C" {

mufft_plan_1d *mufft_create_plan_1d_c2c(unsigned N, int direction, unsigned flags) {
    if (direction != MUFFT_FORWARD && direction != MUFFT_INVERSE) {
        set_err("invalid direction");
        return nullptr;
    }
    return create_1d(KIND_C2C, N, direction, flags, 0);
}

mufft_plan_1d *mufft_create_plan_1d_r2c(unsigned N, unsigned flags) {
    if (!is_pow2(N) || N < 4) {  // N == 2 fails in the reference too (fft.c:485 -> c2c(1))
        set_err("invalid transform size");
        return nullptr;
    }
    return create_1d(KIND_R2C, N / 2, MUFFT_FORWARD, flags, 0);
}

mufft_plan_1d *mufft_create_plan_1d_c2r(unsigned N, unsigned flags) {
    if (!is_pow2(N) || N < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    return create_1d(KIND_C2R, N / 2, MUFFT_INVERSE, flags, 0);
}

mufft_plan_1d *mufft_hip_create_plan_1d_variant(unsigned N, int kind, unsigned flags, int variant) {
    switch (kind) {
    case TK_C2C_FWD:
    case TK_C2C_INV:
        return create_1d(KIND_C2C, N, kind == TK_C2C_FWD ? MUFFT_FORWARD : MUFFT_INVERSE, flags, variant);
    case TK_R2C:
    case TK_C2R:
        if (!is_pow2(N) || N < 4) {
            set_err("invalid transform size");
            return nullptr;
        }
        return kind == TK_R2C ? create_1d(KIND_R2C, N / 2, MUFFT_FORWARD, flags, variant)
                              : create_1d(KIND_C2R, N / 2, MUFFT_INVERSE, flags, variant);
    default:
        set_err("invalid kind");
        return nullptr;
    }
}

int mufft_hip_plan_1d_variant(const mufft_plan_1d *p) {
    if (!p) {
        set_err("null plan");
        return -1;
    }
    return p->eng.row_fused() ? p->eng.row_var : 0;
}

void mufft_execute_plan_1d(mufft_plan_1d *plan, void *MUFFT_RESTRICT output,
                           const void *MUFFT_RESTRICT input) {
    if (!plan) return;
    exec_1d_dropin(plan, output, input, nullptr, 0, 0u, 1.0f);
}

void mufft_free_plan_1d(mufft_plan_1d *plan) { destroy_1d(plan); }

int mufft_hip_execute_plan_1d_batched(mufft_plan_1d *p, void *output, const void *input,
                                      unsigned howmany, size_t idist, size_t odist, void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    long long id, od;
    if (api_dists(p, false, idist, odist, &id, &od)) return -1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const PtrInfo ci = classify_ptr(input), co = classify_ptr(output);
    if (ci.kind == PTR_DEVICE && co.kind == PTR_DEVICE)
        return exec_1d(p, output, input, nullptr, howmany, id, 0, od, 0u, 1.0f, st);
    // Host memory: chunked through device staging, copies overlapping the transforms.
    std::lock_guard<std::mutex> g(p->lock);
    const StagedOperand so{output, (size_t)out_dist_c(p) * 8u, (size_t)od * 8u, co.kind != PTR_DEVICE};
    const StagedOperand si{const_cast<void *>(input), (size_t)in_dist_c(p) * 8u, (size_t)id * 8u,
                           ci.kind != PTR_DEVICE};
    return run_staged(p->pipe, st, so, si, false, howmany, [&](void *o, const void *i, unsigned n) {
        return exec_1d(p, o, i, nullptr, n, id, 0, od, 0u, 1.0f, st);
    });
}

int mufft_hip_execute_adjoint_1d_batched(mufft_plan_1d *p, void *output, const void *input,
                                         unsigned howmany, size_t idist, size_t odist,
                                         void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    long long id, od;
    if (api_dists(p, true, idist, odist, &id, &od) || kernel_ptrs(&output, &input)) return -1;
    return exec_adjoint(p, output, input, nullptr, howmany, id, 0, od, 0u, 1.0f,
                        static_cast<hipStream_t>(stream));
}

int mufft_hip_execute_plan_1d_mul_batched(mufft_plan_1d *p, void *output, const void *input,
                                          const void *input2, unsigned howmany, size_t idist,
                                          size_t idist2, size_t odist, float norm, int conj2,
                                          int adjoint, void *stream) {
    if (!p || !output || !input || !input2) {
        set_err("null argument");
        return -1;
    }
    // The multiply fuses into the prologue of kernels with complex input.
    const int kind_run = adjoint ? (p->kind == KIND_C2C ? KIND_C2C : (p->kind == KIND_R2C ? KIND_C2R : KIND_R2C))
                                 : p->kind;
    if (kind_run == KIND_R2C) {
        set_err("no fused multiply for transforms with real input");
        return -1;
    }
    long long id, od;
    if (api_dists(p, adjoint != 0, idist, odist, &id, &od) || kernel_ptrs(&output, &input, &input2))
        return -1;
    const unsigned extra = RF_MUL | (conj2 ? RF_CONJ2 : 0u);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (adjoint)
        return exec_adjoint(p, output, input, input2, howmany, id, (long long)idist2, od, extra, norm, st);
    return exec_1d(p, output, input, input2, howmany, id, (long long)idist2, od, extra, norm, st);
}

int mufft_hip_plan_1d_launches(const mufft_plan_1d *p) {
    if (!p) return -1;
    if (p->eng.row_fused()) return 1;
    const int c2c = p->eng.four_step() ? (p->eng.three_pass() ? 3 : 2) : p->eng.launches_generic();
    return c2c + (p->kind == KIND_C2C ? 0 : 1);
}

}  // extern "C"
