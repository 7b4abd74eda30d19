// Host-side building blocks: error state, device buffers, pointer
// classification, twiddle tables, the variant table and the Engine.
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "host.h"

namespace mufft {

thread_local char g_err[512] = "";

void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int DevBuf::ensure(size_t n) {
    if (n <= bytes) return 0;
    if (fixed) {
        set_err("scratch of %zu bytes is too small (%zu needed)", bytes, n);
        return -1;
    }
    release();
    if (hipMalloc(&p, n) != hipSuccess) {
        (void)hipGetLastError();
        set_err("hipMalloc(%zu bytes) failed", n);
        p = nullptr;
        return -1;
    }
    bytes = n;
    return 0;
}

void DevBuf::release() {
    if (p && !fixed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void ScratchOrder::release() {
    if (ev) (void)hipEventDestroy(ev);
    ev = nullptr;
    used = false;
    last = nullptr;
}

PtrInfo classify_ptr(const void *ptr) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return {PTR_PAGEABLE, nullptr};
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
        attr.type == hipMemoryTypeArray)
        return {PTR_DEVICE, const_cast<void *>(ptr)};
    if (attr.type == hipMemoryTypeHost && attr.devicePointer) {
        // devicePointer is the mapping of the allocation's base address.
        const char *host_base = static_cast<const char *>(attr.hostPointer ? attr.hostPointer : ptr);
        return {PTR_PINNED, static_cast<char *>(attr.devicePointer) +
                                (static_cast<const char *>(ptr) - host_base)};
    }
    return {PTR_PAGEABLE, nullptr};
}

bool is_host_ptr(const void *ptr) { return classify_ptr(ptr).kind != PTR_DEVICE; }

int kernel_ptr(const void *ptr, const void **out) {
    const PtrInfo pi = classify_ptr(ptr);
    if (pi.kind == PTR_PAGEABLE) {
        set_err("buffer must be device memory or pinned host memory (mufft_alloc)");
        return -1;
    }
    *out = pi.dev;
    return 0;
}

// ---- twiddle tables ----------------------------------------------------------
int build_stockham_table(int n, int direction, const std::vector<int> &radix, float2 **out,
                         hipStream_t s) {
    *out = nullptr;
    if (radix.size() <= 1) return 0;
    if (radix.size() > 24) {
        set_err("radix sequence too long");
        return -1;
    }
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 0;
    a.r0 = radix[0];
    a.npass = (int)radix.size();
    long long p = 1;
    for (size_t i = 0; i < radix.size(); ++i) {
        a.radix[i] = radix[i];
        a.p[i] = p;
        p *= radix[i];
    }
    a.n = n;
    a.count = n - radix[0];
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)a.count));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

int build_resolve_table(int n, int direction, float2 **out, hipStream_t s) {
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 1;
    a.n = n;
    a.count = n;
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)n));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

int build_plain_table(int n, int count, int direction, float2 **out, hipStream_t s) {
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 2;
    a.n = n;
    a.count = count;
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)count));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

int build_split_table(int n, int n2, int direction, float2 **out, hipStream_t s) {
    TwiddleArgs a{};
    a.direction = direction;
    a.mode = 3;
    a.n = n;
    a.p[0] = n2;
    a.count = n - n2;
    HIP_TRY(hipMalloc(out, sizeof(float2) * (size_t)a.count));
    a.out = *out;
    HIP_TRY(launch_twiddles(a, s));
    return 0;
}

std::vector<int> fused_sequence(int n, int E, bool rem_first) {
    if (E > n) E = n;
    const int NP = E >= n ? 1 : (ilog2(n) + ilog2(E) - 1) / ilog2(E);
    std::vector<int> r;
    for (int s = 0; s < NP; ++s) r.push_back(seq_radix(n, E, NP, s, rem_first));
    return r;
}

std::vector<int> generic_sequence(int n) {
    std::vector<int> r;
    int left = n;
    while (left >= 16) {
        r.push_back(16);
        left /= 16;
    }
    if (left > 1) r.push_back(left);
    return r;
}

// ---- variant table -------------------------------------------------------------
namespace {
std::mutex g_var_lock;
int g_variant[ROW_MAX_LOG2 + 1][4] = {};
}  // namespace

int tune_kind(int kind, int dir) {
    if (kind == KIND_C2C) return dir < 0 ? TK_C2C_FWD : TK_C2C_INV;
    return kind == KIND_R2C ? TK_R2C : TK_C2R;
}

int chosen_variant(int logn, int tk) {
    std::lock_guard<std::mutex> g(g_var_lock);
    const int v = g_variant[logn][tk];
    return v ? v : rv_default(1 << logn, tk);
}

void set_chosen_variant(int logn, int tk, int variant) {
    std::lock_guard<std::mutex> g(g_var_lock);
    g_variant[logn][tk] = variant;
}

bool variant_is_set(int logn, int tk) {
    std::lock_guard<std::mutex> g(g_var_lock);
    return g_variant[logn][tk] != 0;
}

namespace {
constexpr bool default_splits_valid() {
    for (int l = ROW_MAX_LOG2 + 1; l <= 2 * COL_MAX_LOG2; ++l)
        if (!fs_code_valid(l, fs_code_default(l))) return false;
    return true;
}
static_assert(default_splits_valid(), "every static split of the four-step sizes is runnable");
int g_fs_code[2 * COL_MAX_LOG2 + 1] = {};
int g_col_tx[COL_MAX_LOG2 + 1] = {};
}  // namespace

int chosen_fs_code(int logn) {
    std::lock_guard<std::mutex> g(g_var_lock);
    const int c = (logn >= 0 && logn <= 2 * COL_MAX_LOG2) ? g_fs_code[logn] : 0;
    return c ? c : fs_code_default(logn);
}

void set_fs_code(int logn, int code) {
    std::lock_guard<std::mutex> g(g_var_lock);
    if (logn >= 0 && logn <= 2 * COL_MAX_LOG2) g_fs_code[logn] = code;
}

int chosen_col_tx(int logL) {
    std::lock_guard<std::mutex> g(g_var_lock);
    const int t = (logL >= 0 && logL <= COL_MAX_LOG2) ? g_col_tx[logL] : 0;
    return t ? t : col_TX(1 << logL);
}

void set_col_tx(int logL, int tx) {
    std::lock_guard<std::mutex> g(g_var_lock);
    if (logL >= 0 && logL <= COL_MAX_LOG2) g_col_tx[logL] = tx;
}

// ---- Engine ----------------------------------------------------------------------
int Engine::init(int n_, int dir, int kind, bool rows, bool cols, hipStream_t s, int var, int fs, int tx) {
    n = n_;
    logn = ilog2(n_);
    direction = dir;
    if (rows && row_fused()) {
        row_var = var ? var : chosen_variant(logn, tune_kind(kind, dir));
        const int n2 = rv_n(n, row_var);
        if (build_stockham_table(n2, dir, fused_sequence(n2, rv_R(n, row_var), true), &tw_row, s)) return -1;
        if (rv_split(row_var) > 0 && build_split_table(n, n2, dir, &tw_split, s)) return -1;
    }
    if (cols && col_fused()) {
        col_tx = tx ? tx : chosen_col_tx(logn);
        if (build_stockham_table(n, dir, fused_sequence(n, col_E(n), false), &tw_col, s)) return -1;
    }
    if (rows && four_step()) {
        fs_code = fs ? fs : chosen_fs_code(logn);
        if (!fs_code_valid(logn, fs_code)) {
            set_err("split %#x is not valid for 2^%d points", fs_code, logn);
            return -1;
        }
        const int f3 = (fs_code & 0xff) ? fs_code : 0;
        log_n1 = fs_code_n1(fs_code);
        log_n2 = logn - log_n1;
        log_s = log_n2;
        const int n1 = 1 << log_n1, n2 = 1 << log_n2;
        if (build_stockham_table(n1, dir, fused_sequence(n1, col_E(n1), false), &tw_fs1, s) ||
            build_plain_table(n >> log_s, n >> log_s, dir, &tw_coarse, s) ||
            build_plain_table(n, 1 << log_s, dir, &tw_fine, s))
            return -1;
        if (f3) {
            log_a = (f3 >> 4) & 15;
            log_b = f3 & 15;
            log_s2 = log_n2 / 2;
            const int na = 1 << log_a, nb = 1 << log_b;
            if (build_stockham_table(na, dir, fused_sequence(na, col_E(na), false), &tw_pa, s) ||
                build_stockham_table(nb, dir, fused_sequence(nb, col_E(nb), false), &tw_pb, s) ||
                build_plain_table(n2 >> log_s2, n2 >> log_s2, dir, &tw_coarse2, s) ||
                build_plain_table(n2, 1 << log_s2, dir, &tw_fine2, s))
                return -1;
        } else if (build_stockham_table(n2, dir, fused_sequence(n2, col_E(n2), false), &tw_fs2, s)) {
            return -1;
        }
    }
    if ((rows && !row_fused() && !four_step()) || (cols && !col_fused())) {
        gen_radix = generic_sequence(n);
        if (build_stockham_table(n, dir, gen_radix, &tw_gen, s)) return -1;
    }
    return 0;
}

void Engine::release() {
    for (float2 *t : {tw_row, tw_col, tw_gen, tw_split, tw_fs1, tw_fs2, tw_coarse, tw_fine, tw_pa, tw_pb,
                      tw_coarse2, tw_fine2})
        if (t) (void)hipFree(t);
    tw_row = tw_col = tw_gen = tw_split = nullptr;
    tw_fs1 = tw_fs2 = tw_coarse = tw_fine = nullptr;
    tw_pa = tw_pb = tw_coarse2 = tw_fine2 = nullptr;
}

}  // namespace mufft
