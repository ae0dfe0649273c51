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
    int* counter;  // zeroed by the launch itself
    const ServeSys* sys;
    float* base;  // arena base
    int grid;
    size_t lds_bytes;
};

hipError_t launch_serve(int method, int mode, const ServeLaunch& p, hipStream_t stream);
hipError_t serve_occupancy(int method, int mode, size_t lds_bytes, int* blocks_per_cu);

}  // namespace tprh
