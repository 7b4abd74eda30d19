// Host-memory batches: the batch runs in chunks through two device staging
// slots on three streams, so the upload of chunk i+1 and the download of
// chunk i-1 overlap the transform of chunk i.  Device layouts mirror the host
// layouts (same distances), so the transform runs unchanged on each chunk.
#pragma once

#include <algorithm>

#include "host.h"

struct HostPipe {
    static constexpr size_t kChunkBytes = 32u << 20;
    hipStream_t up = nullptr, down = nullptr;
    hipEvent_t ev_up[2] = {}, ev_comp[2] = {}, ev_down[2] = {};
    mufft::DevBuf in[2], out[2];
    int init() {
        if (up) return 0;
        HIP_TRY(hipStreamCreateWithFlags(&up, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&down, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            HIP_TRY(hipEventCreateWithFlags(&ev_up[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ev_comp[i], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&ev_down[i], hipEventDisableTiming));
        }
        return 0;
    }
    void release() {
        if (up) (void)hipStreamSynchronize(up);
        if (down) (void)hipStreamSynchronize(down);
        for (int i = 0; i < 2; ++i) {
            if (ev_up[i]) (void)hipEventDestroy(ev_up[i]);
            if (ev_comp[i]) (void)hipEventDestroy(ev_comp[i]);
            if (ev_down[i]) (void)hipEventDestroy(ev_down[i]);
            ev_up[i] = ev_comp[i] = ev_down[i] = nullptr;
            in[i].release();
            out[i].release();
        }
        if (up) (void)hipStreamDestroy(up);
        if (down) (void)hipStreamDestroy(down);
        up = down = nullptr;
    }
};

namespace mufft {

// Pipe of a plan, created on first use (callers hold the plan's lock).
inline HostPipe *plan_pipe(HostPipe *&pp) {
    if (!pp) pp = new (std::nothrow) HostPipe;
    if (!pp) set_err("out of host memory");
    return pp;
}
inline void free_pipe(HostPipe *&pp) {
    if (pp) {
        pp->release();
        delete pp;
        pp = nullptr;
    }
}

// One operand of a staged batch: unit u occupies [u*stride, u*stride + unit)
// bytes from `host`.
struct StagedOperand {
    void *host;
    size_t unit, stride;
    bool staged;
};

// Runs run(dev_out, dev_in, count) over `howmany` units in chunks; returns
// when the last chunk's output is in host memory.  preload_out uploads an
// output operand before the transform (for transforms that leave part of
// their output untouched).
template <typename Run>
int run_staged_chunks(HostPipe &pp, hipStream_t comp, StagedOperand o, StagedOperand i,
                      bool preload_out, unsigned howmany, Run run) {
    const size_t big = std::max(o.stride, i.stride);
    const unsigned per = (unsigned)std::max<size_t>(1, HostPipe::kChunkBytes / std::max<size_t>(big, 1));
    const unsigned nchunks = (howmany + per - 1) / per;
    for (int sl = 0; sl < 2; ++sl) {
        const size_t cnt = std::min<size_t>(per, howmany);
        if (i.staged && pp.in[sl].ensure((cnt - 1) * i.stride + i.unit)) return -1;
        if (o.staged && pp.out[sl].ensure((cnt - 1) * o.stride + o.unit)) return -1;
    }
    // Work queued on `comp` before this call comes first.
    HIP_TRY(hipEventRecord(pp.ev_comp[0], comp));
    HIP_TRY(hipStreamWaitEvent(pp.up, pp.ev_comp[0], 0));
    HIP_TRY(hipStreamWaitEvent(pp.down, pp.ev_comp[0], 0));
    for (unsigned c = 0; c < nchunks; ++c) {
        const int sl = c & 1;
        const unsigned u0 = c * per;
        const unsigned cnt = std::min(per, howmany - u0);
        const size_t in_bytes = (size_t)(cnt - 1) * i.stride + i.unit;
        const size_t out_bytes = (size_t)(cnt - 1) * o.stride + o.unit;
        char *hin = static_cast<char *>(i.host) + (size_t)u0 * i.stride;
        char *hout = static_cast<char *>(o.host) + (size_t)u0 * o.stride;
        const bool up_work = i.staged || (o.staged && preload_out);
        if (c >= 2) {
            // Slot reuse: chunk c-2 must have been transformed and downloaded.
            HIP_TRY(hipStreamWaitEvent(pp.up, pp.ev_comp[sl], 0));
            if (o.staged) HIP_TRY(hipStreamWaitEvent(pp.up, pp.ev_down[sl], 0));
        }
        if (i.staged)
            HIP_TRY(hipMemcpyAsync(pp.in[sl].p, hin, in_bytes, hipMemcpyHostToDevice, pp.up));
        if (o.staged && preload_out)
            HIP_TRY(hipMemcpyAsync(pp.out[sl].p, hout, out_bytes, hipMemcpyHostToDevice, pp.up));
        if (up_work) {
            HIP_TRY(hipEventRecord(pp.ev_up[sl], pp.up));
            HIP_TRY(hipStreamWaitEvent(comp, pp.ev_up[sl], 0));
        }
        if (c >= 2 && o.staged) HIP_TRY(hipStreamWaitEvent(comp, pp.ev_down[sl], 0));
        void *dout = o.staged ? pp.out[sl].p : static_cast<void *>(hout);
        const void *din = i.staged ? pp.in[sl].p : static_cast<const void *>(hin);
        if (run(dout, din, cnt)) return -1;
        HIP_TRY(hipEventRecord(pp.ev_comp[sl], comp));
        if (o.staged) {
            HIP_TRY(hipStreamWaitEvent(pp.down, pp.ev_comp[sl], 0));
            HIP_TRY(hipMemcpyAsync(hout, pp.out[sl].p, out_bytes, hipMemcpyDeviceToHost, pp.down));
            HIP_TRY(hipEventRecord(pp.ev_down[sl], pp.down));
        }
    }
    HIP_TRY(hipStreamSynchronize(pp.down));
    HIP_TRY(hipStreamSynchronize(pp.up));
    HIP_TRY(hipStreamSynchronize(comp));
    return 0;
}

// As run_staged_chunks; after a failure the three streams are drained, so no
// copy into the staging slots is still in flight when the call returns.
template <typename Run>
int run_staged(HostPipe *&pipe, hipStream_t comp, StagedOperand o, StagedOperand i, bool preload_out,
               unsigned howmany, Run run) {
    HostPipe *pp = plan_pipe(pipe);
    if (!pp || pp->init()) return -1;
    const int rc = run_staged_chunks(*pp, comp, o, i, preload_out, howmany, run);
    if (rc) {
        (void)hipStreamSynchronize(pp->up);
        (void)hipStreamSynchronize(pp->down);
        (void)hipStreamSynchronize(comp);
    }
    return rc;
}

}  // namespace mufft
