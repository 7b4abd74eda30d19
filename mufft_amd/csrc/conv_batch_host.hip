// Host side of the uneven convolution batch.
#include <algorithm>
#include <new>

#include "conv_batch.h"
#include "host.h"

using namespace mufft;

// ===========================================================================
// Uneven convolution batch: two persistent launches (forward tasks, then
// inverse tasks) for many convolutions of mixed sizes and methods, replayed
// from a captured hipGraph.
// ===========================================================================
struct mufft_hip_conv_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<CBTask> tasks;
    std::vector<CBJob> jobs;
    CBArgs args{};
    int grid = 0;
    DevBuf d_tasks, d_jobs, d_flags, d_spectra;
    std::vector<float2 *> tables;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void destroy_batch(mufft_hip_conv_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (float2 *t : b->tables)
        if (t) (void)hipFree(t);
    b->d_tasks.release();
    b->d_jobs.release();
    b->d_flags.release();
    b->d_spectra.release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace

extern "C" {

void mufft_hip_free_conv_batch(mufft_hip_conv_batch *b) { destroy_batch(b); }

static mufft_hip_conv_batch *mufft_hip_create_conv_batch_impl(const mufft_hip_conv_job *jobs, unsigned count,
                                                  unsigned flags) {
    (void)flags;  // CPU ISA bits have no meaning here
    if (!jobs || count == 0) {
        set_err("empty job list");
        return nullptr;
    }
    // Validate: mono N in [8, 65536], stereo N in [8, 32768], powers of two.
    for (unsigned i = 0; i < count; ++i) {
        const unsigned N = jobs[i].N;
        const bool stereo = (jobs[i].method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        if (!is_pow2(N) || N < 8 || N > (stereo ? 32768u : 65536u) || !jobs[i].first ||
            !jobs[i].second || !jobs[i].output) {
            set_err("job %u: unsupported size or null buffer", i);
            return nullptr;
        }
    }
    mufft_hip_conv_batch *b = new (std::nothrow) mufft_hip_conv_batch;
    if (!b) return nullptr;
    if (hipGetDevice(&b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_err("no usable HIP device");
        b->stream = nullptr;
        destroy_batch(b);
        return nullptr;
    }
    // Order jobs so equal shapes are adjacent: (method, N) descending.
    std::vector<unsigned> order(count);
    for (unsigned i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](unsigned x, unsigned y) {
        const unsigned sx = jobs[x].method & 1, sy = jobs[y].method & 1;
        if (jobs[x].N != jobs[y].N) return jobs[x].N > jobs[y].N;
        return sx < sy;
    });
    // Spectra: mono (N/2 + 1) bins each; stereo N bins each.
    std::vector<size_t> spec_off(count);
    size_t spec_points = 0;
    for (unsigned r = 0; r < count; ++r) {
        const mufft_hip_conv_job &j = jobs[order[r]];
        const bool stereo = (j.method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        spec_off[r] = spec_points;
        const size_t n = stereo ? j.N : j.N / 2 + 1;
        spec_points += 2 * ((n + 7) & ~(size_t)7);
    }
    if (b->d_spectra.ensure(spec_points * sizeof(float2))) {
        destroy_batch(b);
        return nullptr;
    }
    b->jobs.resize(count);
    for (unsigned r = 0; r < count; ++r) {
        const mufft_hip_conv_job &j = jobs[order[r]];
        const bool stereo = (j.method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        const size_t n = stereo ? j.N : j.N / 2 + 1;
        CBJob &d = b->jobs[r];
        d.first = static_cast<const float2 *>(j.first);
        d.second = static_cast<const float2 *>(j.second);
        d.spec1 = b->d_spectra.f2() + spec_off[r];
        d.spec2 = d.spec1 + ((n + 7) & ~(size_t)7);
        d.out = static_cast<float2 *>(j.output);
    }
    // Twiddle tables per (log2 n, kind, direction) of the persistent kernel's variants.
    // Transforms of 2^CB_SPLIT_LOG2 points use the 2^(CB_SPLIT_LOG2-1)-point
    // sub-transform's table plus the split twiddles.
    struct TabKey {
        int logn, kind, dir;
        float2 *tw, *rtw, *stw;
    };
    std::vector<TabKey> tabs;
    auto table_for = [&](int logn, int kind, int dir, float2 **tw, float2 **rtw, float2 **stw) -> int {
        for (const auto &t : tabs)
            if (t.logn == logn && t.kind == kind && t.dir == dir) {
                *tw = t.tw;
                *rtw = t.rtw;
                *stw = t.stw;
                return 0;
            }
        TabKey k{logn, kind, dir, nullptr, nullptr, nullptr};
        const int n = 1 << logn;
        const int n2 = logn >= CB_SPLIT_LOG2 ? n / 2 : n;
        if (build_stockham_table(n2, dir, fused_sequence(n2, cb_points_per_thread(logn), true), &k.tw, b->stream))
            return -1;
        b->tables.push_back(k.tw);
        if (n2 != n) {
            if (build_split_table(n, n2, dir, &k.stw, b->stream)) return -1;
            b->tables.push_back(k.stw);
        }
        if (kind != KIND_C2C) {
            if (build_resolve_table(n, dir, &k.rtw, b->stream)) return -1;
            b->tables.push_back(k.rtw);
        }
        tabs.push_back(k);
        *tw = k.tw;
        *rtw = k.rtw;
        *stw = k.stw;
        return 0;
    };
    // Tasks: runs of equal (method, N) jobs, split by workgroup capacity.
    std::vector<CBTask> fwd, inv;
    for (unsigned r = 0; r < count;) {
        const mufft_hip_conv_job &j0 = jobs[order[r]];
        const bool stereo = (j0.method & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        unsigned e = r;
        while (e < count && jobs[order[e]].N == j0.N && (jobs[order[e]].method & 1) == (j0.method & 1)) ++e;
        const int lN = ilog2(j0.N);
        // (type, kind, dir, logn, zero-pad, full, mul)
        struct Shape {
            int type, kind, dir, logn;
            unsigned flags;
        };
        std::vector<Shape> shapes;
        if (!stereo) {
            shapes.push_back({CB_FWD_FIRST, KIND_R2C, -1, lN - 1, 0u});
            shapes.push_back({CB_FWD_SECOND, KIND_R2C, -1, lN - 1, 0u});
            shapes.push_back({CB_INV, KIND_C2R, 1, lN - 1, (unsigned)RF_MUL});
        } else {
            shapes.push_back({CB_FWD_FIRST, KIND_C2C, -1, lN, 0u});
            shapes.push_back({CB_FWD_SECOND, KIND_R2C, -1, lN - 1, (unsigned)RF_FULL});
            shapes.push_back({CB_INV, KIND_C2C, 1, lN, (unsigned)RF_MUL});
        }
        for (Shape sh : shapes) {
            // zero padding per operand follows each job's method bits; a task
            // mixes only jobs with equal bits, so split runs on them too.
            const int cap = cb_jobs_per_task(sh.logn);
            for (unsigned q = r; q < e;) {
                const unsigned m0 = jobs[order[q]].method;
                unsigned q2 = q;
                while (q2 < e && q2 - q < (unsigned)cap && jobs[order[q2]].method == m0) ++q2;
                CBTask t{};
                t.type = sh.type;
                t.kind = sh.kind;
                t.dir = sh.dir;
                t.logn = sh.logn;
                t.job0 = (int)q;
                t.njobs = (int)(q2 - q);
                t.flags = sh.flags;
                if (sh.type != CB_INV) t.flags |= RF_KEEP_OUT;   // spectra are read by the inverse phase
                if (sh.type == CB_FWD_FIRST && (m0 & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST))
                    t.flags |= RF_ZERO_PAD;
                if (sh.type == CB_FWD_SECOND && (m0 & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND))
                    t.flags |= RF_ZERO_PAD;
                t.norm = 1.0f / (float)j0.N;
                float2 *tw = nullptr, *rtw = nullptr, *stw = nullptr;
                if (table_for(t.logn, t.kind, t.dir, &tw, &rtw, &stw)) {
                    destroy_batch(b);
                    return nullptr;
                }
                t.tw = tw;
                t.rtw = rtw;
                t.stw = stw;
                t.k1 = 0;
                (sh.type == CB_INV ? inv : fwd).push_back(t);
                if (t.logn >= CB_SPLIT_LOG2) {   // second half of the split transform
                    t.k1 = 1;
                    (sh.type == CB_INV ? inv : fwd).push_back(t);
                }
                q = q2;
            }
        }
        r = e;
    }
    // Longest-processing-time first inside each phase.
    auto cost = [](const CBTask &t) { return (double)t.njobs * (double)(1 << t.logn) * t.logn; };
    std::stable_sort(fwd.begin(), fwd.end(), [&](const CBTask &x, const CBTask &y) { return cost(x) > cost(y); });
    std::stable_sort(inv.begin(), inv.end(), [&](const CBTask &x, const CBTask &y) { return cost(x) > cost(y); });
    b->tasks = fwd;
    b->tasks.insert(b->tasks.end(), inv.begin(), inv.end());

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
        set_err("hipGetDeviceProperties failed");
        destroy_batch(b);
        return nullptr;
    }
    b->grid = std::max(1, std::min<int>(prop.multiProcessorCount, (int)b->tasks.size()));
    if (b->d_tasks.ensure(b->tasks.size() * sizeof(CBTask)) ||
        b->d_jobs.ensure(b->jobs.size() * sizeof(CBJob)) ||
        b->d_flags.ensure(4 * sizeof(int)) ||
        hipMemcpyAsync(b->d_tasks.p, b->tasks.data(), b->tasks.size() * sizeof(CBTask),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_jobs.p, b->jobs.data(), b->jobs.size() * sizeof(CBJob),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        if (!g_err[0]) set_err("batch setup failed");
        destroy_batch(b);
        return nullptr;
    }
    CBArgs &a = b->args;
    a.tasks = static_cast<const CBTask *>(b->d_tasks.p);
    a.nfwd = (int)fwd.size();
    a.ntasks = (int)b->tasks.size();
    a.jobs = static_cast<const CBJob *>(b->d_jobs.p);
    a.njobs = (int)count;
    a.head = static_cast<int *>(b->d_flags.p);
    a.done = a.head + 2;
    // Capture reset + the two persistent launches once; executions replay the graph.
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess ||
        launch_conv_batch(a, b->grid, b->stream) != hipSuccess ||
        hipStreamEndCapture(b->stream, &b->graph) != hipSuccess ||
        hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        set_err("graph capture of the batch failed");
        destroy_batch(b);
        return nullptr;
    }
    return b;
}


mufft_hip_conv_batch *mufft_hip_create_conv_batch(const mufft_hip_conv_job *jobs, unsigned count,
                                                  unsigned flags) {
    return guarded<mufft_hip_conv_batch *>(nullptr, [&] { return mufft_hip_create_conv_batch_impl(jobs, count, flags); });
}

int mufft_hip_execute_conv_batch(mufft_hip_conv_batch *b, void *stream) {
    if (!b) {
        set_err("null batch");
        return -1;
    }
    HIP_TRY(hipGraphLaunch(b->exec, stream ? static_cast<hipStream_t>(stream) : b->stream));
    if (!stream) HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

// Launches the batch without the graph (same kernels, same stream order).
int mufft_hip_execute_conv_batch_direct(mufft_hip_conv_batch *b, void *stream) {
    if (!b) {
        set_err("null batch");
        return -1;
    }
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : b->stream;
    HIP_TRY(launch_conv_batch(b->args, b->grid, s));
    if (!stream) HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// 0 when the last execution ran every task, 1 otherwise.  Synchronises the
// device (the batch may have run on a caller's stream).
int mufft_hip_conv_batch_status(mufft_hip_conv_batch *b) {
    if (!b) return -1;
    int done = 0;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(&done, b->args.done, sizeof(int), hipMemcpyDeviceToHost));
    return done == b->args.ntasks ? 0 : 1;
}

int mufft_hip_conv_batch_info(const mufft_hip_conv_batch *b, int *ntasks, int *grid) {
    if (!b) return -1;
    if (ntasks) *ntasks = (int)b->tasks.size();
    if (grid) *grid = b->grid;
    return 0;
}

}  // extern "C"

