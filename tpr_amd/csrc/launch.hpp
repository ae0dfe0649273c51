// launch.hpp -- host-side interface between the C ABI (capi.hip) and the
// kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

namespace tprh {

// LDS words per system for the slice-boundary exchange and the reduced solve.
__host__ __device__ constexpr int lds_words_per_system(int m) { return 13 * m; }

// Slice shapes (E equations per lane, L lanes per slice) compiled into the
// library, S = E * L.  The default shape for each S is chosen by
// default_shape(); the others are alternatives for the autotuner.
#define TPRH_SLICE_SHAPES(X) \
    X(1, 2)                  \
    X(1, 4)                  \
    X(1, 8)                  \
    X(1, 16)                 \
    X(2, 16)                 \
    X(4, 16)                 \
    X(8, 16)                 \
    X(16, 16)                \
    X(4, 64)                 \
    X(8, 32)                 \
    X(8, 64)                 \
    X(16, 32)                \
    X(16, 64)                \
    X(32, 64)

inline bool shape_compiled(int E, int L) {
#define TPRH_Q(EE, LL) \
    if (E == EE && L == LL) return true;
    TPRH_SLICE_SHAPES(TPRH_Q)
#undef TPRH_Q
    return false;
}

// Default (E, L) for a slice size s; returns false if s is unsupported.
inline bool default_shape(int s, int* E, int* L) {
    switch (s) {
        case 2: *E = 1; *L = 2; return true;
        case 4: *E = 1; *L = 4; return true;
        case 8: *E = 1; *L = 8; return true;
        case 16: *E = 1; *L = 16; return true;
        case 32: *E = 2; *L = 16; return true;
        case 64: *E = 4; *L = 16; return true;
        case 128: *E = 8; *L = 16; return true;
        case 256: *E = 16; *L = 16; return true;
        case 512: *E = 16; *L = 32; return true;
        case 1024: *E = 16; *L = 64; return true;
        case 2048: *E = 32; *L = 64; return true;
        default: return false;
    }
}

constexpr bool is_default_shape(int E, int L) {
    return (E == 1 && (L == 2 || L == 4 || L == 8 || L == 16)) || (L == 16 && (E == 2 || E == 4 || E == 8 || E == 16)) ||
           (E == 16 && (L == 32 || L == 64)) || (E == 32 && L == 64);
}

struct BatchArgs {
    const void* a;  // float, or _Float16 when f16
    const void* b;  // may be null: unit diagonal
    const void* c;
    const float* d;
    float* x;
    int n;     // equations per system
    int m;     // slices per system (n / s)
    int nsys;  // systems
    bool f16 = false;     // a, b, c stored as fp16
    bool refine = false;  // x += A^-1 (d - A x), residual in FP64
    bool shared_coef = false;  // every system uses the one matrix at a, b, c
    bool x_cached = false;     // store x through the caches (false: non-temporal stores)
};

// method 0 = PTPR, 1 = TPR; mode 0 = fast, 1 = exact.
hipError_t launch_batched(int method, int mode, const BatchArgs& p, int E, int L, hipStream_t stream);
hipError_t launch_batched_ptpr(int mode, const BatchArgs& p, int E, int L, hipStream_t stream);
hipError_t launch_batched_tpr(int mode, const BatchArgs& p, int E, int L, hipStream_t stream);

// Raises a kernel's dynamic-LDS limit to `bytes`, once per device (a
// function attribute is per device); thread-safe.  `done` is the caller's
// per-kernel record of the devices already set.
inline hipError_t ensure_lds_limit(const void* fn, int bytes, std::atomic<unsigned long long>& done) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const unsigned long long bit = dev < 64 ? 1ULL << dev : 0ULL;
    if (bit && (done.load(std::memory_order_acquire) & bit)) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess && bit) done.fetch_or(bit, std::memory_order_acq_rel);
    return e;
}

}  // namespace tprh
