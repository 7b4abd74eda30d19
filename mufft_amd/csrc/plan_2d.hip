// 2D plans: creation (fft.c:664-792) and execution (fft.c:852-989).
//
// The reference runs all row passes per row while the row is cache
// resident, then each vertical pass over the whole image; c2r reverses the
// order (fft.c:861-862).  Here a batch of images runs as one fused row launch
// and one fused column launch (tiles of adjacent columns), in the same order.
#include <stdio.h>

#include <new>

#include "host.h"
#include "staging.h"

namespace mufft {

static void destroy_2d(mufft_plan_2d *p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    p->ex.release();
    p->ey.release();
    if (p->rtw) (void)hipFree(p->rtw);
    p->s0.release();
    p->s1.release();
    p->s2.release();
    p->alt.release();
    p->stage_in.release();
    p->stage_out.release();
    free_pipe(p->pipe);
    p->scratch.release();
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static mufft_plan_2d *create_2d_impl(int kind, unsigned Nx_c, unsigned Ny, int direction,
                                     unsigned flags) {
    if (!is_pow2(Nx_c) || !is_pow2(Ny) || Nx_c < 2 || Ny < 2 || Nx_c > (1u << 20) ||
        Ny > (1u << 20)) {
        set_err("invalid transform size");
        return nullptr;
    }
    mufft_plan_2d *p = new (std::nothrow) mufft_plan_2d;
    if (!p) return nullptr;
    p->Nx = Nx_c;
    p->Ny = Ny;
    p->kind = kind;
    p->direction = direction;
    p->full = kind == KIND_R2C && (flags & MUFFT_FLAG_FULL_R2C);
    p->zero_pad = (flags & MUFFT_FLAG_ZERO_PAD_UPPER_HALF) != 0;
    if (kind == KIND_C2C) p->vertical_nx = Nx_c;
    else if (kind == KIND_R2C) p->vertical_nx = p->full ? 2 * Nx_c : Nx_c + 1;
    else p->vertical_nx = Nx_c + 1;
    if (hipGetDevice(&p->device) != hipSuccess ||
        hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_err("no usable HIP device");
        p->stream = nullptr;
        destroy_2d(p);
        return nullptr;
    }
    if (p->ex.init((int)Nx_c, direction, kind, true, false, p->stream) ||
        p->ey.init((int)Ny, direction, KIND_C2C, false, true, p->stream) ||
        (kind != KIND_C2C && build_resolve_table((int)Nx_c, direction, &p->rtw, p->stream))) {
        destroy_2d(p);
        return nullptr;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) {
        set_err("twiddle generation failed");
        destroy_2d(p);
        return nullptr;
    }
    const size_t px = (size_t)Nx_c * Ny;
    if (kind == KIND_C2C) {
        p->in_bytes = px * 8;
        p->out_bytes = px * 8;
    } else if (kind == KIND_R2C) {
        p->in_bytes = px * 8;          // Nx_real * Ny floats
        p->out_bytes = 2 * px * 8;     // Nx_real complex per row
    } else {
        p->in_bytes = 2 * px * 8;
        p->out_bytes = px * 8;         // Nx_real * Ny floats
    }
    return p;
}

static mufft_plan_2d *create_2d(int kind, unsigned Nx_c, unsigned Ny, int direction, unsigned flags) {
    return guarded<mufft_plan_2d *>(nullptr, [&] { return create_2d_impl(kind, Nx_c, Ny, direction, flags); });
}

size_t plan_2d_scratch_points(const mufft_plan_2d *p) {
    const size_t Nx = p->Nx, Ny = p->Ny;
    if (p->kind == KIND_C2R) return (Nx + 1) * Ny;
    if (!p->ey.col_fused()) return (p->kind == KIND_C2C ? Nx : 2 * Nx) * Ny;
    return 0;
}

static int exec_2d_images(mufft_plan_2d *p, void *out_, const void *in_, unsigned howmany, DevBuf &s0,
                          DevBuf &s1, DevBuf &s2, DevBuf &alt, hipStream_t st) {
    const long long Nx = p->Nx, Ny = p->Ny;
    const int hm = (int)howmany;
    const float2 *in = static_cast<const float2 *>(in_);
    float2 *out = static_cast<float2 *>(out_);
    const unsigned zp = p->zero_pad ? RF_ZERO_PAD : 0u;
    if (p->kind == KIND_C2C || p->kind == KIND_R2C) {
        const long long rs = p->kind == KIND_C2C ? Nx : 2 * Nx;   // output row stride
        const long long img = rs * Ny;
        float2 *rows_dst = out;
        if (!p->ey.col_fused()) {
            if (s2.ensure(sizeof(float2) * (size_t)img * hm)) return -1;
            rows_dst = s2.f2();
        }
        // The column pass reads the rows' output next: keep it cacheable.
        const unsigned f = zp | (p->full ? RF_FULL : 0u) | RF_KEEP_OUT;
        if (run_rows(p->ex, p->kind, p->rtw, in, Nx, nullptr, 0, rows_dst, rs, (int)(Ny * hm), f,
                     1.0f, s0, s1, st))
            return -1;
        return run_cols(p->ey, rows_dst, rs, img, out, rs, img, (int)p->vertical_nx, hm, alt, st);
    }
    // c2r: columns first (fft.c:863-916), then the per-row resolve and rows.
    const long long in_rs = 2 * Nx, in_img = in_rs * Ny;
    const long long mid_rs = Nx + 1, mid_img = mid_rs * Ny;
    if (s2.ensure(sizeof(float2) * (size_t)mid_img * hm)) return -1;
    if (run_cols(p->ey, in, in_rs, in_img, s2.f2(), mid_rs, mid_img, (int)(Nx + 1), hm, alt, st))
        return -1;
    return run_rows(p->ex, KIND_C2R, p->rtw, s2.f2(), mid_rs, nullptr, 0, out, Nx, (int)(Ny * hm),
                    zp, 1.0f, s0, s1, st);
}

// A c2c batch of images of 8-32 MiB (1024^2 to 2048^2 points) that holds
// more than this many bytes runs in balanced chunks of at most this size,
// each chunk's row pass and column pass back to back: the rows' output then
// stays in the 256 MB Infinity Cache for the column pass.  Measured on
// MI355X, 512 MiB per batch: c2c 1024^2 x 64 445 -> 389 us, 2048^2 x 16
// 583 -> 454 us; but c2c 512^2 x 256 +2%, r2c / c2r 2048^2 x 16 +6-7% and
// r2c 1024^2 x 64 +4% (their padded rows leave half of the output buffer
// untouched), so those run whole.
constexpr size_t k2dChunkBytes = size_t(256) << 20;

int exec_2d_with(mufft_plan_2d *p, void *out, const void *in, unsigned howmany, DevBuf &s0,
                 DevBuf &s1, DevBuf &s2, DevBuf &alt, hipStream_t st) {
    const size_t img = p->in_bytes;
    const size_t per = k2dChunkBytes / img;
    if (p->kind != KIND_C2C || img < (size_t(8) << 20) || img > (size_t(32) << 20) || howmany <= per)
        return exec_2d_images(p, out, in, howmany, s0, s1, s2, alt, st);
    const size_t nchunks = (howmany + per - 1) / per;
    const unsigned ch = (unsigned)((howmany + nchunks - 1) / nchunks);
    for (unsigned i = 0; i < howmany; i += ch) {
        const unsigned n = howmany - i < ch ? howmany - i : ch;
        if (exec_2d_images(p, static_cast<char *>(out) + (size_t)i * p->out_bytes,
                           static_cast<const char *>(in) + (size_t)i * p->in_bytes, n, s0, s1, s2, alt, st))
            return -1;
    }
    return 0;
}

// howmany packed images with the plan's own scratch.
static int exec_2d(mufft_plan_2d *p, void *out, const void *in, unsigned howmany, hipStream_t st) {
    auto run = [&] { return exec_2d_with(p, out, in, howmany, p->s0, p->s1, p->s2, p->alt, st); };
    if (plan_2d_scratch_points(p) == 0 && !rows_use_scratch(p->ex)) return run();
    return with_scratch(p->scratch, st, run);
}

}  // namespace mufft

using namespace mufft;

extern "C" {

mufft_plan_2d *mufft_create_plan_2d_c2c(unsigned Nx, unsigned Ny, int direction, unsigned flags) {
    if (direction != MUFFT_FORWARD && direction != MUFFT_INVERSE) {
        set_err("invalid direction");
        return nullptr;
    }
    return create_2d(KIND_C2C, Nx, Ny, direction, flags);
}

mufft_plan_2d *mufft_create_plan_2d_r2c(unsigned Nx, unsigned Ny, unsigned flags) {
    if (!is_pow2(Nx) || Nx < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    return create_2d(KIND_R2C, Nx / 2, Ny, MUFFT_FORWARD, flags);
}

mufft_plan_2d *mufft_create_plan_2d_c2r(unsigned Nx, unsigned Ny, unsigned flags) {
    if (!is_pow2(Nx) || Nx < 4) {
        set_err("invalid transform size");
        return nullptr;
    }
    return create_2d(KIND_C2R, Nx / 2, Ny, MUFFT_INVERSE, flags);
}

void mufft_execute_plan_2d(mufft_plan_2d *p, void *MUFFT_RESTRICT output,
                           const void *MUFFT_RESTRICT input) {
    if (!p) return;
    std::lock_guard<std::mutex> g(p->lock);
    int dev_prev = 0;
    (void)hipGetDevice(&dev_prev);
    if (dev_prev != p->device) (void)hipSetDevice(p->device);
    const void *din = input;
    void *dout = output;
    const bool hin = is_host_ptr(input), hout = is_host_ptr(output);
    int rc = 0;
    if (hin) {
        rc |= p->stage_in.ensure(p->in_bytes);
        if (!rc) rc |= hipMemcpyAsync(p->stage_in.p, input, p->in_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
        din = p->stage_in.p;
    }
    if (!rc && hout) {
        rc |= p->stage_out.ensure(p->out_bytes);
        dout = p->stage_out.p;
        // Columns the transform leaves untouched come back as they went in.
        if (!rc) rc |= hipMemcpyAsync(p->stage_out.p, output, p->out_bytes, hipMemcpyHostToDevice, p->stream) != hipSuccess;
    }
    if (!rc) rc = exec_2d(p, dout, din, 1, p->stream);
    if (!rc && hout)
        rc |= hipMemcpyAsync(output, p->stage_out.p, p->out_bytes, hipMemcpyDeviceToHost, p->stream) != hipSuccess;
    if (hipStreamSynchronize(p->stream) != hipSuccess || rc) {
        if (!rc) set_err("execution failed: %s", hipGetErrorString(hipGetLastError()));
        fprintf(stderr, "mufft_hip: %s\n", g_err);
    }
    if (dev_prev != p->device) (void)hipSetDevice(dev_prev);
}

void mufft_free_plan_2d(mufft_plan_2d *plan) { destroy_2d(plan); }

int mufft_hip_execute_plan_2d_batched(mufft_plan_2d *p, void *output, const void *input,
                                      unsigned howmany, void *stream) {
    if (!p || !output || !input) {
        set_err("null argument");
        return -1;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const PtrInfo ci = classify_ptr(input), co = classify_ptr(output);
    if (ci.kind == PTR_DEVICE && co.kind == PTR_DEVICE) return exec_2d(p, output, input, howmany, st);
    // Host memory: chunked through device staging, copies overlapping the
    // transforms.  A padded r2c output keeps the columns the transform does
    // not write, so those images are uploaded first.
    std::lock_guard<std::mutex> g(p->lock);
    const bool preload = p->kind == KIND_R2C && !p->full;
    const StagedOperand so{output, p->out_bytes, p->out_bytes, co.kind != PTR_DEVICE};
    const StagedOperand si{const_cast<void *>(input), p->in_bytes, p->in_bytes, ci.kind != PTR_DEVICE};
    return run_staged(p->pipe, st, so, si, preload, howmany, [&](void *o, const void *i, unsigned n) {
        return exec_2d(p, o, i, n, st);
    });
}

int mufft_hip_plan_2d_launches(const mufft_plan_2d *p) {
    if (!p) return -1;
    int rows = p->ex.row_fused() ? 1
             : (p->ex.four_step() ? (p->ex.three_pass() ? 3 : 2) : p->ex.launches_generic()) + (p->kind == KIND_C2C ? 0 : 1);
    int cols = p->ey.col_fused() ? 1 : p->ey.launches_generic();
    return rows + cols;
}

}  // extern "C"
