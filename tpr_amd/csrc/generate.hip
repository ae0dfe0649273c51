// generate.hip -- seeded counter-based generator of the synthetic systems.
//
// Every value is a pure function of (seed, system id, stream, index), so a
// system's bits never depend on batch composition, and the host copy of the
// generator (in the test oracle) produces byte-identical arrays.  Built with
// -ffp-contract=off like the rest of the library, so the fp32 arithmetic of
// b and the fp64 accumulation of d = A * xt round exactly as on the host.
//
//   key(seed, sys)             = mix64(seed + G * (sys + 1)),   G = 0x9E3779B97F4A7C15
//   draw(key, stream, i)       = mix64(key + G * (8 * i + stream + 1)) >> 40   (24 bits, k)
//   U(-1,1) = (k - 2^23) * 2^-23,   U[0,1) = k * 2^-24          (both exact in fp32)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch_gen.hpp"

namespace tprh {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

__device__ __forceinline__ uint32_t draw24(uint64_t key, uint64_t stream, uint64_t i) {
    return (uint32_t)(mix64(key + kGolden * (8 * i + stream + 1)) >> 40);
}
__device__ __forceinline__ float u_pm1(uint32_t k) { return (float)((int32_t)k - 8388608) * (1.0f / 8388608.0f); }
__device__ __forceinline__ float u_01(uint32_t k) { return (float)k * (1.0f / 16777216.0f); }

enum { kStA = 0, kStC = 1, kStB = 2, kStXt = 3, kStRhs = 4, kStMulti = 5 };

__device__ __forceinline__ void gen_coef(uint64_t key, int i, int n, int kind, float& a, float& b, float& c) {
    float ai = u_pm1(draw24(key, kStA, (uint64_t)i));
    float ci = u_pm1(draw24(key, kStC, (uint64_t)i));
    if (i == 0) ai = 0.0f;
    if (i == n - 1) ci = 0.0f;
    float u = u_01(draw24(key, kStB, (uint64_t)i));
    float extra = kind == 0 ? (1.0f + u) : (1e-4f + u * 0.0099f);
    a = ai;
    c = ci;
    b = (fabsf(ai) + fabsf(ci)) + extra;
}

// Equation i of system `sys` (0 <= i < n) at array position g.
__device__ __forceinline__ void generate_one(int kind, uint64_t key, int i, int n, size_t g, float* __restrict__ a,
                                             float* __restrict__ b, float* __restrict__ c, float* __restrict__ d,
                                             float* __restrict__ xt) {
    if (kind == 2 || kind == 3) {
        float ai = (float)(-1.0 / 6.0), ci = (float)(-1.0 / 6.0);
        if (i == 0) ai = 0.0f;
        if (i == n - 1) ci = 0.0f;
        a[g] = ai;
        c[g] = ci;
        b[g] = 1.0f;
        d[g] = kind == 2 ? (float)(1.0 * (i + 1)) : u_pm1(draw24(key, kStRhs, (uint64_t)i));
        return;
    }
    float ai, bi, ci;
    gen_coef(key, i, n, kind, ai, bi, ci);
    const float xi = u_pm1(draw24(key, kStXt, (uint64_t)i));
    double acc = (double)bi * (double)xi;
    if (i > 0) {
        const float xl = u_pm1(draw24(key, kStXt, (uint64_t)(i - 1)));
        acc = (double)ai * (double)xl + acc;
    }
    if (i < n - 1) {
        const float xr = u_pm1(draw24(key, kStXt, (uint64_t)(i + 1)));
        acc = acc + (double)ci * (double)xr;
    }
    a[g] = ai;
    b[g] = bi;
    c[g] = ci;
    d[g] = (float)acc;
    if (xt) xt[g] = xi;
}

__global__ void generate_kernel(int kind, uint64_t seed, uint64_t sys0, int nsys, int n, float* __restrict__ a,
                                float* __restrict__ b, float* __restrict__ c, float* __restrict__ d,
                                float* __restrict__ xt) {
    const size_t total = (size_t)nsys * (size_t)n;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(g / (size_t)n);
        const int i = (int)(g - (size_t)k * n);
        const uint64_t key = mix64(seed + kGolden * (sys0 + (uint64_t)k + 1));
        generate_one(kind, key, i, n, g, a, b, c, d, xt);
    }
}

// One workgroup per system; system k has n[k] equations at offsets[k].
__global__ void generate_uneven_kernel(int kind, uint64_t seed, uint64_t sys0, int nsys, const int* __restrict__ n,
                                       const long long* __restrict__ offsets, float* __restrict__ a,
                                       float* __restrict__ b, float* __restrict__ c, float* __restrict__ d,
                                       float* __restrict__ xt) {
    for (int k = blockIdx.x; k < nsys; k += gridDim.x) {
        const int nk = n[k];
        const size_t base = (size_t)offsets[k];
        const uint64_t key = mix64(seed + kGolden * (sys0 + (uint64_t)k + 1));
        for (int i = threadIdx.x; i < nk; i += blockDim.x) generate_one(kind, key, i, nk, base + i, a, b, c, d, xt);
    }
}

// Right-hand sides j = j0 .. j0+nrhs-1 of system `sys`: d ~ U(-1,1), written
// back to back (nrhs * n floats).
__global__ void generate_rhs_kernel(uint64_t seed, uint64_t sys, int j0, int nrhs, int n, float* __restrict__ d) {
    const size_t total = (size_t)nrhs * (size_t)n;
    const uint64_t key = mix64(seed + kGolden * (sys + 1));
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const uint64_t jj = (uint64_t)j0 + g / (size_t)n;
        const uint64_t i = g % (size_t)n;
        d[g] = u_pm1(draw24(key, kStMulti, jj * (uint64_t)n + i));
    }
}

// Equation i of the coefficients of system `sys` only, packed [a | b | c]
// (n floats each): the same values generate_one writes.
__device__ __forceinline__ void coef_one(int kind, uint64_t key, int i, int n, float* __restrict__ abc) {
    float ai, bi, ci;
    if (kind == 2 || kind == 3) {
        ai = i == 0 ? 0.0f : (float)(-1.0 / 6.0);
        ci = i == n - 1 ? 0.0f : (float)(-1.0 / 6.0);
        bi = 1.0f;
    } else {
        gen_coef(key, i, n, kind, ai, bi, ci);
    }
    abc[i] = ai;
    abc[n + i] = bi;
    abc[2 * (size_t)n + i] = ci;
}

// A flat grid over the concatenated values of every job, kGenPerThread
// values per thread at a stride of the block size (coalesced stores).  The
// block finds the job of its first value by binary search over the jobs'
// first indices; every thread then walks forward through the jobs as its
// values advance (a block's 2048 values span a few jobs at most), keeping
// the job's fields and key in registers.
constexpr int kGenBlock = 256;
constexpr int kGenPerThread = 8;

__global__ __launch_bounds__(kGenBlock) void generate_jobs_kernel(const GenJob* __restrict__ jobs, int njobs,
                                                                 long long total, float* __restrict__ base) {
    const long long start = (long long)blockIdx.x * (kGenBlock * kGenPerThread);
    int lo = 0, hi = njobs - 1;  // the last job with first <= start
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= start) lo = mid;
        else hi = mid - 1;
    }
    int j = lo;
    GenJob jb = jobs[j];
    uint64_t key = mix64(jb.seed + kGolden * (jb.sys + 1));
    long long next = j + 1 < njobs ? jobs[j + 1].first : total;
#pragma unroll
    for (int i = 0; i < kGenPerThread; i++) {
        const long long g = start + (long long)i * kGenBlock + threadIdx.x;
        if (g >= total) break;
        if (g >= next) {
            while (j + 1 < njobs && jobs[j + 1].first <= g) j++;
            jb = jobs[j];
            key = mix64(jb.seed + kGolden * (jb.sys + 1));
            next = j + 1 < njobs ? jobs[j + 1].first : total;
        }
        const uint64_t e = (uint64_t)(g - jb.first);  // value index inside the job
        float* dst = base + jb.dst;
        if (jb.coef_kind == 0) {
            const uint64_t jj = (uint64_t)jb.j0 + (e >> jb.lg_n);
            const uint64_t idx = (e & ((1ULL << jb.lg_n) - 1));
            dst[e] = u_pm1(draw24(key, kStMulti, (jj << jb.lg_n) + idx));
        } else {
            coef_one(jb.coef_kind - 1, key, (int)e, 1 << jb.lg_n, dst);
        }
    }
}

hipError_t launch_generate_jobs(const GenJob* jobs, int njobs, long long total, float* base, hipStream_t stream) {
    if (njobs == 0 || total == 0) return hipSuccess;
    const long long per_block = (long long)kGenBlock * kGenPerThread;
    const long long blocks = (total + per_block - 1) / per_block;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(generate_jobs_kernel, dim3((unsigned)blocks), dim3(kGenBlock), 0, stream, jobs, njobs, total,
                       base);
    return hipGetLastError();
}

hipError_t launch_generate(int kind, uint64_t seed, uint64_t sys0, int nsys, int n, float* a, float* b, float* c,
                           float* d, float* xt, hipStream_t stream) {
    const size_t total = (size_t)nsys * (size_t)n;
    size_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, kind, seed, sys0, nsys, n, a,
                       b, c, d, xt);
    return hipGetLastError();
}

hipError_t launch_generate_uneven(int kind, uint64_t seed, uint64_t sys0, int nsys, const int* n,
                                  const long long* offsets, float* a, float* b, float* c, float* d, float* xt,
                                  hipStream_t stream) {
    const int blocks = nsys < 65536 ? nsys : 65536;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(generate_uneven_kernel, dim3(blocks), dim3(256), 0, stream, kind, seed, sys0, nsys, n, offsets, a,
                       b, c, d, xt);
    return hipGetLastError();
}

hipError_t launch_generate_rhs(uint64_t seed, uint64_t sys, int j0, int nrhs, int n, float* d, hipStream_t stream) {
    const size_t total = (size_t)nrhs * (size_t)n;
    size_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(generate_rhs_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, seed, sys, j0, nrhs, n, d);
    return hipGetLastError();
}

}  // namespace tprh
