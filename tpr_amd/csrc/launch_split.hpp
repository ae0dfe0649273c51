// launch_split.hpp -- host-side interface of the multi-workgroup (split) path.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace tprh {

// Slice shapes compiled for the split path (the default shape of every s).
#define TPRH_SPLIT_SHAPES(X) \
    X(1, 2)                  \
    X(1, 4)                  \
    X(1, 8)                  \
    X(1, 16)                 \
    X(2, 16)                 \
    X(4, 16)                 \
    X(8, 16)                 \
    X(16, 16)                \
    X(16, 32)                \
    X(16, 64)                \
    X(32, 64)

// Largest reduced system (slices per system) the split path solves in one
// workgroup's LDS: 10 rows of m floats.
constexpr int kSplitMaxM = 2048;

struct SplitArgs {
    const void* a;  // float or _Float16 (f16)
    const void* b;  // may be null: unit diagonal
    const void* c;
    const float* d;
    float* x;
    int n, nsys, chunk;
    bool f16, refine;
    bool shared_coef;  // every system uses the one matrix at a, b, c (several right-hand sides)
    int form = 0;      // 0 = the library's choice, 1 = two passes, 2 = one pass
    // workspace, for one round of `chunk` systems
    float* rows;  // chunk * m * kSplitRowWords: boundary rows (and x) per slice
    float* xed;   // chunk * m: x at the end of every slice
    int* sync;    // kSplitSyncStride * (chunk + 1) words: the one-pass form's counters and flags
};

// Floats per slice in the boundary buffer (see split.hip).
constexpr int kSplitRowWords = 8;

// Synchronisation words per system of the one-pass form: one 128-byte line.
constexpr int kSplitSyncStride = 32;

// Bytes of workspace for rounds of `chunk` systems of m slices: the rows and
// x_ed per slice, and the synchronisation words of the one-pass form.
inline size_t split_workspace_bytes(int m, int chunk) {
    return sizeof(float) * (size_t)chunk * (kSplitRowWords + 1) * (size_t)m +
           sizeof(int) * kSplitSyncStride * ((size_t)chunk + 1) + 256;
}

// Carves the workspace (256-byte aligned) into the SplitArgs pointers.
inline void split_carve(void* ws, int m, int chunk, SplitArgs* p) {
    size_t addr = ((size_t)ws + 255) & ~(size_t)255;
    float* f = reinterpret_cast<float*>(addr);
    p->rows = f;
    f += (size_t)chunk * m * kSplitRowWords;
    p->xed = f;
    f += (size_t)chunk * m;
    p->sync = reinterpret_cast<int*>(f);
}

hipError_t launch_split(int method, int mode, const SplitArgs& p, int E, int L, hipStream_t stream);

}  // namespace tprh
