// launch_uneven.hpp -- host-side interface of the persistent uneven-batch
// kernel (uneven.hip) and its plan (capi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace tprh {

// Slice shapes the persistent kernel carries a code path for: (id, E, L).
// Its 1024-lane workgroups leave 128 VGPRs per lane, so the menu stops at
// the shapes whose fast path fits (E = 16 with L > 16 and E = 32 need more);
// every size class up to n = 16384 has a shape here (s = 256, E = L = 16).
#define TPRH_UNEVEN_SHAPES(X) \
    X(0, 1, 16)               \
    X(1, 2, 16)               \
    X(2, 4, 16)               \
    X(3, 8, 16)               \
    X(4, 16, 16)              \
    X(5, 8, 32)               \
    X(6, 4, 64)               \
    X(7, 8, 64)

inline int uneven_shape_id(int E, int L) {
#define TPRH_UNEVEN_ID(ID, EE, LL) \
    if (E == EE && L == LL) return ID;
    TPRH_UNEVEN_SHAPES(TPRH_UNEVEN_ID)
#undef TPRH_UNEVEN_ID
    return -1;
}

// One work item: `count` systems of one size class, as many as one
// 1024-lane workgroup holds (1024 / (m * L)).
struct UnevenTask {
    int shape;     // TPRH_UNEVEN_SHAPES id
    int m;         // slices per system
    int lg_lanes;  // log2(m * L): lanes per system
    int first;     // first entry of the plan's system list
    int count;     // systems in the item
};

struct UnevenLaunch {
    const UnevenTask* tasks;
    int ntasks;
    int* counter;
    const int* sys_list;
    const long long* offsets;
    const float *a, *b, *c, *d;
    float* x;
    int grid;
    size_t lds_bytes;
};

hipError_t launch_uneven(int method, int mode, const UnevenLaunch& p, hipStream_t stream);
hipError_t uneven_occupancy(int method, int mode, size_t lds_bytes, int* blocks_per_cu);

}  // namespace tprh
