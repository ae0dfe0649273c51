// exact_div.hip -- a diagnostic build (libtpr_diag.so, not part of the
// solver library): checks the solver's short reciprocal (recip_rn in
// tpr_amd/csrc/tpr_device.hpp) against IEEE division, 1.0f / x, for every
// one of the 2^32 float bit patterns x on the GPU.  Equal means the same
// bits, or both NaN.  Counts go to counts[0] (recip_rn, the guarded form the
// kernels use), counts[1] (the unguarded four-instruction sequence) and,
// for the unguarded one, counts[2] (x subnormal), counts[3]
// (2^-126 <= |x| < 2^-125 or |x| > 2^126) and counts[4] (the rest); up to 64
// patterns where the guarded form differs go to bad[].
#include <hip/hip_runtime.h>

#include "../csrc/tpr_device.hpp"

namespace {

__device__ bool same(float a, float b) {
    return __builtin_bit_cast(unsigned, a) == __builtin_bit_cast(unsigned, b) || (a != a && b != b);
}

__global__ void recip_check_kernel(unsigned long long* __restrict__ counts, unsigned* __restrict__ bad) {
    const unsigned long long total = 1ull << 32;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const unsigned u = (unsigned)i;
        const float x = __builtin_bit_cast(float, u);
        const float ref = 1.0f / x;
        if (!same(tprh::recip_rn(x), ref)) {
            c0++;
            const unsigned long long k = atomicAdd(&counts[5], 1ull);
            if (k < 64) bad[k] = u;
        }
        if (!same(tprh::recip_rn_unguarded(x), ref)) {
            c1++;
            const unsigned ax = u & 0x7fffffffu;
            if (ax < 0x00800000u) c2++;
            else if (ax < 0x01000000u || ax > 0x7e800000u) c3++;
            else c4++;
        }
    }
    if (c0) atomicAdd(&counts[0], c0);
    if (c1) atomicAdd(&counts[1], c1);
    if (c2) atomicAdd(&counts[2], c2);
    if (c3) atomicAdd(&counts[3], c3);
    if (c4) atomicAdd(&counts[4], c4);
}

}  // namespace

// counts: 6 zeroed words (the last one counts recorded patterns); bad: 64
// words.  Returns a hipError_t.
extern "C" int tprh_diag_recip_check(unsigned long long* counts, unsigned* bad, void* stream) {
    hipLaunchKernelGGL(recip_check_kernel, dim3(8192), dim3(256), 0, (hipStream_t)stream, counts, bad);
    return (int)hipGetLastError();
}
