// launch_serve.hpp -- host-side interface of the serving batch kernel
// (serve.hip) used by the request loop (server.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "launch_uneven.hpp"

namespace tprh {

// One system of a serving batch: right-hand side j of a request.  Offsets
// are in floats from the server's arena base.  The matrix is stored packed,
// [a | b | c] with n floats each, and is shared by all right-hand sides of
// the request (and by every request on that matrix).
struct ServeSys {
    long long coef;  // a at coef, b at coef + n, c at coef + 2n
    long long rhs;   // d
    long long x;     // answer
};

struct ServeLaunch {
    const UnevenTask* tasks;  // work items (the shape ids of launch_uneven.hpp)
    int ntasks;
    int* counter;  // work counter: must hold 0 when the launch starts (the
                   // serving loop's table copy writes it, no launch of its own)
    const ServeSys* sys;
    float* base;  // arena base
    int grid;
    size_t lds_bytes;
};

// A system of a serving batch whose class needs the multi-workgroup path:
// offsets in floats from the arena base; rows / xed are the system's part of
// its request's workspace (m * 8 boundary floats, m slice-end values).
struct ServeSplitSys {
    long long coef;  // packed a | b | c, n floats each
    long long rhs;
    long long x;
    long long rows;
    long long xed;
};

struct ServeSplitLaunch {
    const ServeSplitSys* sys;  // device table, nsys entries
    int nsys;
    int n;  // every system of one launch has the same size class
    float* base;
};

// Slice shapes the serving batch's split launches carry (s >= 256, the
// default shape of each s).
inline bool serve_split_shape(int E, int L) {
    return (E == 16 && (L == 16 || L == 32 || L == 64)) || (E == 32 && L == 64);
}

hipError_t launch_serve(int method, int mode, const ServeLaunch& p, hipStream_t stream);
// The three split kernels over every system of the table (one size class).
hipError_t launch_serve_split(int method, int mode, const ServeSplitLaunch& p, int E, int L, hipStream_t stream);
hipError_t serve_occupancy(int method, int mode, size_t lds_bytes, int* blocks_per_cu);

}  // namespace tprh
