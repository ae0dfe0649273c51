// serve_split.hip -- the serving loop's launches for size classes above one
// workgroup: the split path's three kernels (split_kernels.hpp) over a table
// of systems, each with its own matrix, right-hand side, answer and
// workspace in the server's arena, so that every such system of a batch --
// all right-hand sides of all its large requests -- shares one set of
// launches.
#include <hip/hip_runtime.h>

#include "launch_serve.hpp"
#include "split_kernels.hpp"

namespace tprh {

struct TableLoc {
    const ServeSplitSys* __restrict__ sys;
    float* base;
    int n, m;
    __device__ const float* A(long long q) const { return base + sys[q].coef; }
    __device__ const float* B(long long q) const { return base + sys[q].coef + n; }
    __device__ const float* C(long long q) const { return base + sys[q].coef + 2 * (long long)n; }
    __device__ const float* D(long long q) const { return base + sys[q].rhs; }
    __device__ float* X(long long q) const { return base + sys[q].x; }
    __device__ float* Rows(long long q) const { return base + sys[q].rows; }
    __device__ float* Xed(long long q) const { return base + sys[q].xed; }
};

namespace {

template <int METHOD, int E, int L, int MODE>
hipError_t launch_el(const ServeSplitLaunch& p, hipStream_t stream) {
    const int m = p.n / (E * L);
    const int lg_m = __builtin_ctz((unsigned)m);
    static std::atomic<unsigned long long> lds_set{0};  // K2 keeps 10 rows of m floats in LDS
    hipError_t le = ensure_lds_limit(reinterpret_cast<const void*>(reduced_kernel<METHOD, TableLoc>),
                                     (int)(sizeof(float) * 10 * kSplitMaxM), lds_set);
    if (le != hipSuccess) return le;
    TableLoc loc{p.sys, p.base, p.n, m};
    const int threads = 256;
    const long long nslices = (long long)p.nsys * m;
    const int grid = (int)((nslices + threads / L - 1) / (threads / L));
    int rthreads = m < 1024 ? m : 1024;
    rthreads = (rthreads + 63) / 64 * 64;
    const size_t rshmem = sizeof(float) * 10 * (size_t)m;
    hipLaunchKernelGGL((rows_kernel<METHOD, E, L, MODE, true, float, false, TableLoc>), dim3(grid), dim3(threads), 0, stream,
                       loc, lg_m, nslices);
    hipLaunchKernelGGL((reduced_kernel<METHOD, TableLoc>), dim3(p.nsys), dim3(rthreads), rshmem, stream, loc);
    hipLaunchKernelGGL((stage3_kernel<METHOD, E, L, MODE, true, float, false, TableLoc>), dim3(grid), dim3(threads), 0,
                       stream, loc, lg_m, nslices);
    return hipGetLastError();
}

template <int METHOD, int MODE>
hipError_t launch_mode(const ServeSplitLaunch& p, int E, int L, hipStream_t stream) {
    if (E == 16 && L == 16) return launch_el<METHOD, 16, 16, MODE>(p, stream);
    if (E == 16 && L == 32) return launch_el<METHOD, 16, 32, MODE>(p, stream);
    if (E == 16 && L == 64) return launch_el<METHOD, 16, 64, MODE>(p, stream);
    if (E == 32 && L == 64) return launch_el<METHOD, 32, 64, MODE>(p, stream);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_serve_split(int method, int mode, const ServeSplitLaunch& p, int E, int L, hipStream_t stream) {
    if (p.nsys == 0) return hipSuccess;
    if (method == 0) return mode == kExact ? launch_mode<0, kExact>(p, E, L, stream) : launch_mode<0, kFast>(p, E, L, stream);
    return mode == kExact ? launch_mode<1, kExact>(p, E, L, stream) : launch_mode<1, kFast>(p, E, L, stream);
}

}  // namespace tprh
