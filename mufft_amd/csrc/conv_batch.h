// Task/job tables of the persistent uneven convolution batch (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include "geometry.h"

namespace mufft {

constexpr int CB_WG = 1024;           // threads per workgroup of the persistent kernel
constexpr int CB_LDS_BYTES = 139264;  // 1024 threads x 17 padded points x 8 bytes
constexpr int CB_MIN_LOG2 = 2;        // smallest fused transform inside the kernel (4 points)
constexpr int CB_SPLIT_LOG2 = 15;     // 32768 points: two workgroups per transform

enum CBTaskType : int { CB_FWD_FIRST = 0, CB_FWD_SECOND = 1, CB_INV = 2 };

// One convolution job: device pointers of its inputs, spectra and output.
struct CBJob {
    const float2 *first;   // first input (N/2 reals or N/2 complex when zero padded)
    const float2 *second;  // second input (real)
    float2 *spec1, *spec2; // spectra of the two inputs (scratch owned by the batch)
    float2 *out;           // N reals (mono) or N complex (stereo)
};

// A group of up to CB_WG / T jobs sharing one transform shape.
struct CBTask {
    int type;              // CBTaskType
    int kind;              // KIND_C2C / KIND_R2C / KIND_C2R
    int dir;               // -1 forward, +1 inverse
    int logn;              // log2 of the fused (complex) transform length
    int job0, njobs;       // jobs [job0, job0 + njobs) of the job table
    unsigned flags;        // RowFlag bits
    float norm;            // 1/N for the inverse multiply
    const float2 *tw;      // Stockham twiddles of the transform
    const float2 *rtw;     // real-transform resolve twiddles (r2c / c2r)
    // Transforms of 2^CB_SPLIT_LOG2 points run as two workgroup tasks
    // (k1 = 0, 1) of the split transform (split_impl.h); stw is its twiddle table.
    int k1;
    const float2 *stw;
};

// Tasks [0, nfwd) are the forward transforms, [nfwd, ntasks) the inverse
// transforms; each phase is one persistent launch with its own queue head.
struct CBArgs {
    const CBTask *tasks;
    int nfwd, ntasks;
    const CBJob *jobs;
    int njobs;
    int *head;             // queue heads of the two phases
    int *done;             // tasks completed (both phases)
};

hipError_t launch_conv_batch(const CBArgs &a, int grid, hipStream_t s);
int cb_jobs_per_task(int logn);
int cb_points_per_thread(int logn);   // of the (sub-)transform a task runs

}  // namespace mufft
