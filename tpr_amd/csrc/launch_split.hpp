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

// Tickets of the pipelined form (pipelined_split_kernel).  A round of nsys
// systems of bps = 2^lg_bps blocks has three kinds of task: rows tasks
// (block i, i < nsys bps), reduced tasks (system q) and stage-3 tasks
// (block i).  Stream A is the rows tasks in order.  Stream B is the reduced
// tasks of systems 0 .. g-1, then for each system b >= g its reduced task
// followed by the stage-3 tasks of system b - g, then the stage-3 tasks of
// the last g systems.  Tickets 0 .. D-1 go to A; then A and B alternate;
// then B to its end.  Every task's dependencies (rows tasks of q before the
// reduced task of q before the stage-3 tasks of q) have lower tickets for
// any D >= 2 bps.
enum { kPipeRows = 0, kPipeReduced = 1, kPipeStage3 = 2 };

struct PipeTask {
    int kind;  // kPipeRows, kPipeReduced, kPipeStage3
    int i;     // block index (rows, stage 3), or q * bps (reduced task of system q)
};

// D: the lag in rows tasks, at least 2 bps and at most the round's rows tasks.
__host__ __device__ inline int pipeline_d(int lag, int na, int bps) {
    const int d = lag > 2 * bps ? lag : 2 * bps;
    return d < na ? d : na;
}

// g: systems between a reduced task and the stage-3 tasks after it.
__host__ __device__ inline int pipeline_gap(int d, int bps, int nsys) {
    const int g = d / (2 * bps);
    return g < 1 ? 1 : (g > nsys ? nsys : g);
}

// Tickets of a round: nsys bps rows tasks, nsys reduced, nsys bps stage 3.
__host__ __device__ inline int pipeline_tickets(int nsys, int lg_bps) { return nsys * ((2 << lg_bps) + 1); }

// The task of ticket tk (0 <= tk < pipeline_tickets), for D and g as above.
__host__ __device__ inline PipeTask pipeline_task(int tk, int nsys, int lg_bps, int D, int g) {
    const int bps = 1 << lg_bps;
    const int pairs = (nsys << lg_bps) - D;
    int k = -1;  // index in stream B
    PipeTask t{kPipeRows, tk};
    if (tk >= D) {
        if (tk < D + 2 * pairs) {
            const int u = tk - D;
            if (u & 1) k = u >> 1;
            else t.i = D + (u >> 1);
        } else {
            k = pairs + (tk - D - 2 * pairs);
        }
    }
    if (k >= 0) {
        const int full = (nsys - g) * (bps + 1);
        if (k < g) {
            t.kind = kPipeReduced;
            t.i = k << lg_bps;
        } else if (k - g < full) {
            const int b = (k - g) / (bps + 1), r = (k - g) - b * (bps + 1);
            t.kind = r == 0 ? kPipeReduced : kPipeStage3;
            t.i = r == 0 ? (b + g) << lg_bps : (b << lg_bps) + r - 1;
        } else {
            t.kind = kPipeStage3;
            t.i = ((nsys - g) << lg_bps) + (k - g - full);
        }
    }
    return t;
}

hipError_t launch_split(int method, int mode, const SplitArgs& p, int E, int L, hipStream_t stream);

}  // namespace tprh
