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

// A table of generation jobs in one launch.  A job with coef_kind == 0
// writes right-hand sides j0 .. j0+nrhs-1 of system `sys` (as
// generate_rhs_kernel) to base + dst, nrhs * n values; a job with
// coef_kind == k + 1 writes the coefficients of system `sys` for generator
// kind k (the a, b, c that generate_kernel writes), packed [a | b | c] at base + dst, and
// counts n values (nrhs is 1).  The launch covers the concatenation of
// every job's values; `first` is the job's starting index in it (jobs in
// increasing order).
struct GenJob {
    uint64_t seed, sys;
    long long dst;    // floats from base
    long long first;  // index of the job's first value in the launch
    int j0, nrhs;
    int lg_n;         // n = 2^lg_n
    int coef_kind;    // 0: right-hand sides; k + 1: coefficients of kind k
};
hipError_t launch_generate_jobs(const GenJob* jobs, int njobs, long long total, float* base, hipStream_t stream);

// Largest n the single-workgroup PCR / CR kernels accept.
constexpr int kSingleMaxN = 4096;

// Batched whole-system PCR (method 2) or CR (method 3): one workgroup per
// system, nsys systems of n equations back to back; b == nullptr means a unit
// diagonal.
hipError_t launch_single(int method, const float* a, const float* b, const float* c, const float* r, float* x, int n,
                         int nsys, hipStream_t stream);

}  // namespace tprh
