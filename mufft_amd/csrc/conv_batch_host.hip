// Host side of the uneven convolution batch.
#include <algorithm>
#include <new>

#include "conv_batch.h"
#include "host.h"

using namespace mufft;

// ===========================================================================
// Uneven convolution batch: one persistent launch (forward tasks, a
// device-wide barrier, inverse tasks) for many convolutions of mixed sizes
// and methods, replayed from a captured hipGraph.
// ===========================================================================
struct mufft_hip_conv_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<CBTask> tasks;
    std::vector<CBJob> jobs;
    CBArgs args{};
    int grid = 0;
    unsigned long long executions = 0;   // launches issued (graph or direct)
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
    // Order jobs so equal shapes are adjacent: N descending, then method.
    std::vector<unsigned> order(count);
    for (unsigned i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](unsigned x, unsigned y) {
        if (jobs[x].N != jobs[y].N) return jobs[x].N > jobs[y].N;
        return jobs[x].method < jobs[y].method;
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
    // Job groups: runs of equal (N, method) jobs, chunked by the largest
    // workgroup capacity of the three transforms; each transform of a group
    // runs as tasks of its own capacity (two halves each at 2^CB_SPLIT_LOG2
    // points).  The group's inverse tasks wait for all its forward tasks.
    struct Group {
        std::vector<CBTask> fwd, inv;
        double cost = 0;
    };
    std::vector<Group> groups;
    for (unsigned r = 0; r < count;) {
        const mufft_hip_conv_job &j0 = jobs[order[r]];
        const unsigned m0 = j0.method;
        const bool stereo = (m0 & 1) == MUFFT_CONV_METHOD_FLAG_STEREO_MONO;
        unsigned e = r;
        while (e < count && jobs[order[e]].N == j0.N && jobs[order[e]].method == m0) ++e;
        const int lN = ilog2(j0.N);
        struct Shape {
            int type, kind, dir, logn;
            unsigned flags;
        };
        Shape shapes[3];
        if (!stereo) {
            shapes[0] = {CB_FWD_FIRST, KIND_R2C, -1, lN - 1, 0u};
            shapes[1] = {CB_FWD_SECOND, KIND_R2C, -1, lN - 1, 0u};
            shapes[2] = {CB_INV, KIND_C2R, 1, lN - 1, (unsigned)RF_MUL};
        } else {
            shapes[0] = {CB_FWD_FIRST, KIND_C2C, -1, lN, 0u};
            shapes[1] = {CB_FWD_SECOND, KIND_R2C, -1, lN - 1, (unsigned)RF_FULL};
            shapes[2] = {CB_INV, KIND_C2C, 1, lN, (unsigned)RF_MUL};
        }
        int gcap = 1;
        for (const Shape &sh : shapes) gcap = std::max(gcap, cb_jobs_per_task(sh.logn));
        for (unsigned q = r; q < e; q += (unsigned)gcap) {
            const unsigned gend = std::min<unsigned>(e, q + (unsigned)gcap);
            Group g;
            const int gi = (int)groups.size();
            int nprod = 0;
            for (const Shape &sh : shapes) {
                const unsigned cap = (unsigned)cb_jobs_per_task(sh.logn);
                float2 *tw = nullptr, *rtw = nullptr, *stw = nullptr;
                if (table_for(sh.logn, sh.kind, sh.dir, &tw, &rtw, &stw)) {
                    destroy_batch(b);
                    return nullptr;
                }
                for (unsigned p = q; p < gend; p += cap) {
                    CBTask t{};
                    t.type = sh.type;
                    t.kind = sh.kind;
                    t.dir = sh.dir;
                    t.logn = sh.logn;
                    t.job0 = (int)p;
                    t.njobs = (int)std::min<unsigned>(cap, gend - p);
                    t.flags = sh.flags;
                    if (sh.type != CB_INV) t.flags |= RF_KEEP_OUT;   // spectra are read by the inverse tasks
                    if (sh.type == CB_FWD_FIRST && (m0 & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_FIRST))
                        t.flags |= RF_ZERO_PAD;
                    if (sh.type == CB_FWD_SECOND && (m0 & MUFFT_CONV_METHOD_FLAG_ZERO_PAD_UPPER_HALF_SECOND))
                        t.flags |= RF_ZERO_PAD;
                    t.norm = 1.0f / (float)j0.N;
                    t.tw = tw;
                    t.rtw = rtw;
                    t.stw = stw;
                    t.group = gi;
                    const int halves = t.logn >= CB_SPLIT_LOG2 ? 2 : 1;
                    for (int k1 = 0; k1 < halves; ++k1) {
                        t.k1 = k1;
                        (sh.type == CB_INV ? g.inv : g.fwd).push_back(t);
                        g.cost += (double)t.njobs * (double)(1 << (t.logn - (halves - 1))) * t.logn;
                    }
                    if (sh.type != CB_INV) nprod += halves;
                }
            }
            for (CBTask &t : g.inv) t.nprod = nprod;
            groups.push_back(std::move(g));
        }
        r = e;
    }
    // Groups to XCDs, longest first onto the least loaded; within an XCD the
    // forward tasks longest first, then the inverse tasks in the same order.
    std::vector<int> gorder(groups.size());
    for (size_t i = 0; i < groups.size(); ++i) gorder[i] = (int)i;
    std::stable_sort(gorder.begin(), gorder.end(),
                     [&](int x, int y) { return groups[x].cost > groups[y].cost; });
    double load[CB_XCDS] = {};
    std::vector<int> gx(groups.size());
    for (int gi : gorder) {
        int best = 0;
        for (int xx = 1; xx < CB_XCDS; ++xx)
            if (load[xx] < load[best]) best = xx;
        gx[gi] = best;
        load[best] += groups[gi].cost;
    }
    CBArgs &a = b->args;
    for (int xx = 0; xx < CB_XCDS; ++xx) {
        a.qbegin[xx] = (int)b->tasks.size();
        for (int gi : gorder)
            if (gx[gi] == xx) b->tasks.insert(b->tasks.end(), groups[gi].fwd.begin(), groups[gi].fwd.end());
        for (int gi : gorder)
            if (gx[gi] == xx) b->tasks.insert(b->tasks.end(), groups[gi].inv.begin(), groups[gi].inv.end());
    }
    a.qbegin[CB_XCDS] = (int)b->tasks.size();
    a.ngroups = (int)groups.size();

    // Every XCD needs workgroups of its own: launch all the resident ones
    // (the dispatcher deals workgroups to the XCDs in turn).
    const int resident = cb_resident_workgroups();
    if (resident < CB_XCDS) {
        set_err("occupancy query of the batch kernel failed");
        destroy_batch(b);
        return nullptr;
    }
    b->grid = resident / CB_XCDS * CB_XCDS;
    // head[CB_XCDS] | ready[ngroups] as unsigned; finish, done, timeouts as 64-bit
    const size_t n32 = (size_t)CB_XCDS + groups.size();
    const size_t off64 = (n32 * sizeof(unsigned) + 7) & ~(size_t)7;
    const size_t flag_bytes = off64 + 3 * sizeof(unsigned long long);
    if (b->d_tasks.ensure(b->tasks.size() * sizeof(CBTask)) ||
        b->d_jobs.ensure(b->jobs.size() * sizeof(CBJob)) ||
        b->d_flags.ensure(flag_bytes) ||
        hipMemcpyAsync(b->d_tasks.p, b->tasks.data(), b->tasks.size() * sizeof(CBTask),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->d_jobs.p, b->jobs.data(), b->jobs.size() * sizeof(CBJob),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemsetAsync(b->d_flags.p, 0, flag_bytes, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        if (!g_err[0]) set_err("batch setup failed");
        destroy_batch(b);
        return nullptr;
    }
    a.tasks = static_cast<const CBTask *>(b->d_tasks.p);
    a.jobs = static_cast<const CBJob *>(b->d_jobs.p);
    a.njobs = (int)count;
    a.head = static_cast<unsigned *>(b->d_flags.p);
    a.ready = a.head + CB_XCDS;
    a.finish = reinterpret_cast<unsigned long long *>(static_cast<char *>(b->d_flags.p) + off64);
    a.done = a.finish + 1;
    a.timeouts = a.finish + 2;
    a.spin_limit = 1u << 24;   // about two seconds of polling
    // Capture the launch once; executions replay the graph.  The counters
    // need no reset between launches (see CBArgs).
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
    ++b->executions;
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
    ++b->executions;
    if (!stream) HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// 0 when every execution so far ran every task behind a completed barrier,
// 1 otherwise.  Synchronises the device (the batch may have run on a
// caller's stream).
int mufft_hip_conv_batch_status(mufft_hip_conv_batch *b) {
    if (!b) return -1;
    unsigned long long done = 0, timeouts = 0;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(&done, b->args.done, sizeof(done), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&timeouts, b->args.timeouts, sizeof(timeouts), hipMemcpyDeviceToHost));
    return done == b->executions * (unsigned long long)b->tasks.size() && timeouts == 0 ? 0 : 1;
}

int mufft_hip_conv_batch_info(const mufft_hip_conv_batch *b, int *ntasks, int *grid) {
    if (!b) return -1;
    if (ntasks) *ntasks = (int)b->tasks.size();
    if (grid) *grid = b->grid;
    return 0;
}

}  // extern "C"

