// single.hip -- whole-system PCR and CR in one workgroup each: the
// counterparts of the reference's test solvers PTPR_CU::pcr_cu and
// TPR_CU::cr_cu (src/ptpr.cu:432-500,683-707; src/tpr.cu:428-540,697-723),
// batched over systems.  The system sits in LDS; the arithmetic is the exact
// operation order of the CPU classes PCR (src/pcr.cpp:15-89) and CR
// (src/cr.cpp:23-116).
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "launch_gen.hpp"
#include "tpr_device.hpp"

namespace tprh {

// Rows of one system into LDS; with a diagonal b, each row is divided by it
// as PCR::set_tridiagonal_system(a, diag, c, rhs) does (src/pcr.hpp:42-53).
__device__ __forceinline__ void load_rows(const float* __restrict__ a, const float* __restrict__ b,
                                          const float* __restrict__ c, const float* __restrict__ r, size_t base,
                                          int n, float* la, float* lc, float* lr) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        float va = a[base + i], vc = c[base + i], vr = r[base + i];
        if (b != nullptr) {
            float vb = b[base + i];
            va = va / vb;
            vc = vc / vb;
            vr = vr / vb;
        }
        la[i] = va;
        lc[i] = vc;
        lr[i] = vr;
    }
}

__global__ __launch_bounds__(1024) void pcr_single_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                          const float* __restrict__ c, const float* __restrict__ r,
                                                          float* __restrict__ x, int n) {
    extern __shared__ float lds[];
    const size_t base = (size_t)blockIdx.x * n;
    float* la = lds;
    float* lc = la + n;
    float* lr = lc + n;
    load_rows(a, b, c, r, base, n, la, lc, lr);
    __syncthreads();
    const float* res = lds_pcr(la, lc, lr, lr + n, lr + 2 * n, lr + 3 * n, n, threadIdx.x, blockDim.x);
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[base + i] = res[i];
}

__global__ __launch_bounds__(1024) void cr_single_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                         const float* __restrict__ c, const float* __restrict__ r,
                                                         float* __restrict__ x, int n) {
    extern __shared__ float lds[];
    const size_t base = (size_t)blockIdx.x * n;
    float* la = lds;
    float* lc = la + n;
    float* lr = lc + n;
    float* lx = lr + n;
    load_rows(a, b, c, r, base, n, la, lc, lr);
    __syncthreads();
    lds_cr(la, lc, lr, lx, n, threadIdx.x, blockDim.x);
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[base + i] = lx[i];
}

hipError_t launch_single(int method, const float* a, const float* b, const float* c, const float* r, float* x, int n,
                         int nsys, hipStream_t stream) {
    // PCR keeps 6 rows of n floats in LDS: 96 KiB at n = 4096, above the
    // 64 KiB default of dynamic LDS.
    static std::atomic<unsigned long long> lds_set{0};
    hipError_t le = ensure_lds_limit(reinterpret_cast<const void*>(pcr_single_kernel),
                                     (int)(sizeof(float) * 6 * kSingleMaxN), lds_set);
    if (le != hipSuccess) return le;
    int threads = n < 1024 ? n : 1024;
    threads = (threads + 63) / 64 * 64;
    if (method == 2) {
        hipLaunchKernelGGL(pcr_single_kernel, dim3(nsys), dim3(threads), sizeof(float) * 6 * (size_t)n, stream, a, b,
                           c, r, x, n);
    } else {
        hipLaunchKernelGGL(cr_single_kernel, dim3(nsys), dim3(threads), sizeof(float) * 4 * (size_t)n, stream, a, b,
                           c, r, x, n);
    }
    return hipGetLastError();
}

}  // namespace tprh
