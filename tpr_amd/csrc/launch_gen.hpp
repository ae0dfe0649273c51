// launch_gen.hpp -- host-side launchers of the generator and the single-system
// PCR / CR kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tprh {

hipError_t launch_generate(int kind, uint64_t seed, uint64_t sys0, int nsys, int n, float* a, float* b, float* c,
                           float* d, float* xt, hipStream_t stream);
hipError_t launch_generate_uneven(int kind, uint64_t seed, uint64_t sys0, int nsys, const int* n,
                                  const long long* offsets, float* a, float* b, float* c, float* d, float* xt,
                                  hipStream_t stream);
hipError_t launch_generate_rhs(uint64_t seed, uint64_t sys, int j0, int nrhs, int n, float* d, hipStream_t stream);
hipError_t launch_generate_coef(int kind, uint64_t seed, uint64_t sys, int n, float* abc, hipStream_t stream);

// Right-hand sides j0 .. j0+nrhs-1 of system `sys` (as generate_rhs_kernel)
// written to base + dst, for a table of requests in one launch.  The
// launch covers the concatenation of every job's nrhs * n values; `first`
// is the job's starting index in it (jobs in increasing order).
struct RhsJob {
    uint64_t seed, sys;
    long long dst;    // floats from base
    long long first;  // index of the job's first value in the launch
    int j0, nrhs;
    int lg_n;         // n = 2^lg_n
    int pad;
};
hipError_t launch_generate_rhs_jobs(const RhsJob* jobs, int njobs, long long total, float* base,
                                    hipStream_t stream);

// Largest n the single-workgroup PCR / CR kernels accept.
constexpr int kSingleMaxN = 4096;

// Batched whole-system PCR (method 2) or CR (method 3): one workgroup per
// system, nsys systems of n equations back to back; b == nullptr means a unit
// diagonal.
hipError_t launch_single(int method, const float* a, const float* b, const float* c, const float* r, float* x, int n,
                         int nsys, hipStream_t stream);

}  // namespace tprh
