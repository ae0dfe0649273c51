// uneven.hip -- batches of systems of different sizes in one persistent launch.
//
// A plan (built on the host, see capi.hip) sorts the systems into size
// classes, gives every class a slice size s and slice shape (E, L) from a
// per-class table (the autotuner's), and cuts each class into work items of
// as many systems as one 1024-lane workgroup holds.  The persistent kernel
// keeps one workgroup per resident slot; each workgroup takes the next work
// item from an atomic counter (items are ordered largest systems first) and
// solves it with the shape's code path, until the items run out; it asks
// for its next ticket as it starts an item, so the counter's latency hides
// under the item's loads.  Every
// system's result depends only on its own size class's (s, E, L), never on
// the batch around it.
#include <hip/hip_runtime.h>

#include "launch_uneven.hpp"
#include "slice_solvers.hpp"
#include "solve_slot.hpp"

namespace tprh {

template <int METHOD, int MODE, bool HAS_B, int E, int L>
__device__ __forceinline__ void run_task(const UnevenTask& tk, const int* __restrict__ sys_list,
                                         const long long* __restrict__ offsets, const float* __restrict__ a,
                                         const float* __restrict__ b, const float* __restrict__ c,
                                         const float* __restrict__ d, float* __restrict__ x, float* lds) {
    const int slot = threadIdx.x >> tk.lg_lanes;
    const bool valid = slot < tk.count;
    const int sys = sys_list[tk.first + (valid ? slot : 0)];
    const size_t x0 = (size_t)offsets[sys];
    solve_slot<METHOD, MODE, HAS_B, E, L>(tk.m, tk.lg_lanes, valid, a + x0, HAS_B ? b + x0 : nullptr, c + x0, d + x0,
                                          x + x0, lds);
}

__global__ void zero_counter_kernel(int* counter) { *counter = 0; }

template <int METHOD, int MODE, bool HAS_B>
__global__ __launch_bounds__(1024) void uneven_kernel(const UnevenTask* __restrict__ tasks, int ntasks,
                                                      int* __restrict__ counter, const int* __restrict__ sys_list,
                                                      const long long* __restrict__ offsets,
                                                      const float* __restrict__ a, const float* __restrict__ b,
                                                      const float* __restrict__ c, const float* __restrict__ d,
                                                      float* __restrict__ x) {
    extern __shared__ float lds[];
    __shared__ int item;
    if (threadIdx.x == 0) item = atomicAdd(counter, 1);
    __syncthreads();
    int it = item;
    while (it < ntasks) {  // uniform: every workgroup leaves once the items are gone
        // the next item's ticket is asked for now and read after this item:
        // the counter is device-coherent, its answer takes microseconds
        // behind the memory traffic, and this item's loads now cover that
        // wait (a workgroup holds at most two tickets; items are independent)
        int nxt = ntasks;
        if (threadIdx.x == 0) nxt = atomicAdd(counter, 1);
        const UnevenTask tk = tasks[it];
        switch (tk.shape) {
#define TPRH_UNEVEN_CASE(ID, EE, LL)                                                                     \
    case ID:                                                                                             \
        run_task<METHOD, MODE, HAS_B, EE, LL>(tk, sys_list, offsets, a, b, c, d, x, lds);                \
        break;
            TPRH_UNEVEN_SHAPES(TPRH_UNEVEN_CASE)
#undef TPRH_UNEVEN_CASE
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

template <int METHOD, int MODE, bool HAS_B>
hipError_t launch_uneven_t(const UnevenLaunch& p, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, p.counter);
    hipLaunchKernelGGL((uneven_kernel<METHOD, MODE, HAS_B>), dim3(p.grid), dim3(1024), p.lds_bytes, stream, p.tasks,
                       p.ntasks, p.counter, p.sys_list, p.offsets, p.a, p.b, p.c, p.d, p.x);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_uneven(int method, int mode, const UnevenLaunch& p, hipStream_t stream) {
    const bool hb = p.b != nullptr;
    if (method == 0) {
        if (mode == kExact) return hb ? launch_uneven_t<0, kExact, true>(p, stream) : launch_uneven_t<0, kExact, false>(p, stream);
        return hb ? launch_uneven_t<0, kFast, true>(p, stream) : launch_uneven_t<0, kFast, false>(p, stream);
    }
    if (mode == kExact) return hb ? launch_uneven_t<1, kExact, true>(p, stream) : launch_uneven_t<1, kExact, false>(p, stream);
    return hb ? launch_uneven_t<1, kFast, true>(p, stream) : launch_uneven_t<1, kFast, false>(p, stream);
}

hipError_t uneven_occupancy(int method, int mode, size_t lds_bytes, int* blocks_per_cu) {
    const void* f;
    if (method == 0) f = mode == kExact ? reinterpret_cast<const void*>(uneven_kernel<0, kExact, true>)
                                        : reinterpret_cast<const void*>(uneven_kernel<0, kFast, true>);
    else f = mode == kExact ? reinterpret_cast<const void*>(uneven_kernel<1, kExact, true>)
                            : reinterpret_cast<const void*>(uneven_kernel<1, kFast, true>);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 1024, lds_bytes);
}

}  // namespace tprh
