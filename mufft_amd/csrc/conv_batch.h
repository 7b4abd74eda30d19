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

// Dies (XCDs) of the GPU; each has its own L2.
constexpr int CB_XCDS = 8;

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
    int group;             // readiness counter of the task's job group
    int nprod;             // inverse tasks: forward tasks the group needs first
};

// One persistent launch runs the whole batch.  Each job group (the jobs of
// one task shape) is assigned to one XCD, and that XCD's workgroups (read
// from the hardware XCC_ID register) pull its tasks from the XCD's own
// queue: first the forward tasks, then the inverse tasks.  A forward task
// waits until its stores are complete (they are then in the XCD's L2) and
// counts itself into the group's readiness counter; an inverse task waits for
// the count, invalidates its CU's L1 and reads the spectra from the shared
// L2.  A producer and its consumer never sit on different XCDs, so no
// device-wide cache write-back is needed (the agent-scope release that would
// order them across XCDs, buffer_wbl2, costs about as much as a task on
// gfx950).  An inverse task only waits for forward tasks dequeued before it
// by running workgroups, so the waits always end; they are bounded all the
// same.  The last workgroup to finish clears the queue and readiness
// counters for the next launch.
struct CBArgs {
    const CBTask *tasks;
    int qbegin[CB_XCDS + 1];      // tasks of XCD x: [qbegin[x], qbegin[x + 1])
    const CBJob *jobs;
    int njobs;
    int ngroups;
    unsigned *head;               // ticket counter per XCD
    unsigned *ready;              // per job group: forward tasks completed
    unsigned long long *finish;   // workgroups finished (all launches)
    unsigned long long *done;     // tasks completed (all launches)
    unsigned long long *timeouts; // readiness waits that gave up (the batch then failed)
    unsigned spin_limit;          // polls of a readiness counter before giving up
};

hipError_t launch_conv_batch(const CBArgs &a, int grid, hipStream_t s);
// Workgroups of the persistent kernel that stay resident at once.
int cb_resident_workgroups();
int cb_jobs_per_task(int logn);
int cb_points_per_thread(int logn);   // of the (sub-)transform a task runs

}  // namespace mufft
