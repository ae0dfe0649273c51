// serve.hip -- the serving loop's batch kernel: systems of different sizes,
// several right-hand sides per matrix, all in one persistent launch.
//
// Same scheduling as the uneven-batch kernel (work items of one size class,
// taken from an atomic counter, largest systems first); each system entry
// names its own matrix, right-hand side and answer in the server's arena, so
// every right-hand side of a request reads the request's one resident copy
// of a, b, c (HBM traffic 12 B per matrix equation + 8 B per right-hand-side
// equation when the coefficients stay in cache between its systems).
#include <hip/hip_runtime.h>

#include "launch_serve.hpp"
#include "solve_slot.hpp"

namespace tprh {

template <int METHOD, int MODE>
__global__ __launch_bounds__(1024) void serve_kernel(const UnevenTask* __restrict__ tasks, int ntasks,
                                                     int* __restrict__ counter, const ServeSys* __restrict__ sys,
                                                     float* __restrict__ base) {
    extern __shared__ float lds[];
    __shared__ int item;
    if (threadIdx.x == 0) item = atomicAdd(counter, 1);
    __syncthreads();
    int it = item;
    while (it < ntasks) {  // uniform: every workgroup leaves once the items are gone
        // the next ticket is asked for now and read after this item, so the
        // device-coherent counter's latency hides under the item's loads
        int nxt = ntasks;
        if (threadIdx.x == 0) nxt = atomicAdd(counter, 1);
        const UnevenTask tk = tasks[it];
        const int slot = threadIdx.x >> tk.lg_lanes;
        const bool valid = slot < tk.count;
        const ServeSys e = sys[tk.first + (valid ? slot : 0)];
        switch (tk.shape) {
#define TPRH_SERVE_CASE(ID, EE, LL)                                                                        \
    case ID: {                                                                                             \
        const size_t n = (size_t)tk.m * (EE * LL);                                                         \
        const float* pa = base + e.coef;                                                                   \
        solve_slot<METHOD, MODE, true, EE, LL>(tk.m, tk.lg_lanes, valid, pa, pa + n, pa + 2 * n, base + e.rhs, \
                                               base + e.x, lds);                                           \
        break;                                                                                             \
    }
            TPRH_UNEVEN_SHAPES(TPRH_SERVE_CASE)
#undef TPRH_SERVE_CASE
            default:
                break;
        }
        __syncthreads();  // the item's LDS rows are read before the next item writes them;
                          // every thread has read `item` before it is rewritten
        if (threadIdx.x == 0) item = nxt;
        __syncthreads();
        it = item;
    }
}

namespace {

template <int METHOD, int MODE>
hipError_t launch_serve_t(const ServeLaunch& p, hipStream_t stream) {
    hipLaunchKernelGGL((serve_kernel<METHOD, MODE>), dim3(p.grid), dim3(1024), p.lds_bytes, stream, p.tasks, p.ntasks,
                       p.counter, p.sys, p.base);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_serve(int method, int mode, const ServeLaunch& p, hipStream_t stream) {
    if (method == 0) return mode == kExact ? launch_serve_t<0, kExact>(p, stream) : launch_serve_t<0, kFast>(p, stream);
    return mode == kExact ? launch_serve_t<1, kExact>(p, stream) : launch_serve_t<1, kFast>(p, stream);
}

hipError_t serve_occupancy(int method, int mode, size_t lds_bytes, int* blocks_per_cu) {
    const void* f;
    if (method == 0) f = mode == kExact ? reinterpret_cast<const void*>(serve_kernel<0, kExact>)
                                        : reinterpret_cast<const void*>(serve_kernel<0, kFast>);
    else f = mode == kExact ? reinterpret_cast<const void*>(serve_kernel<1, kExact>)
                            : reinterpret_cast<const void*>(serve_kernel<1, kFast>);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 1024, lds_bytes);
}

}  // namespace tprh
