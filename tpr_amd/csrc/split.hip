// split.hip -- PTPR / TPR for systems larger than one workgroup.
//
// The reference needs every block of a system co-resident for its grid-wide
// barriers and limits n to s^2 (src/ptpr.cu:543,581,655).  Here a system of m
// slices is solved by three ordinary launches:
//   K1 rows_kernel   : every lane group reduces its slices and writes two
//                      boundary rows per slice (E_st before its update, E_ed);
//   K2 reduced_kernel: one workgroup per system applies E_st += s1 E_ed
//                      (src/ptpr.cpp:193-211), builds the m reduced rows
//                      (:218-241) and solves them in LDS (PCR for PTPR,
//                      src/pcr.cpp; CR for TPR, src/cr.cpp);
//   K3 stage3_kernel : every lane group reloads its slices, redoes stage 1
//                      and runs stage 3 (src/ptpr.cpp:263-287 /
//                      src/tpr.cpp:277-308).
// K1 needs only the first and last row of every slice.  In PTPR's in-slice
// PCR the final value of a slice's first (last) row depends only on the rows
// at offsets 0 mod 2^p (-1 mod 2^p) of the previous levels, updated with the
// same formulas in the same order as TPR's tree stage 1 (src/tpr.cpp:142-204
// vs src/ptpr.cpp:140-182), so K1 runs the O(S) tree for both methods and its
// rows are bit-identical to the full PCR's.  Nothing but those rows is
// stored between the kernels: HBM traffic is 16 B/eq (K1) + 16 B/eq + 4 B/eq
// (K3), with the stage-1 work of K1 a quarter of a full PTPR stage 1.
#include <hip/hip_runtime.h>

#include "split_kernels.hpp"

namespace tprh {

namespace {

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
void launch_round(const SplitArgs& p, size_t base, int grid, int threads, int lg_m, long long nslices, int cs, int m,
                  int rthreads, size_t rshmem, hipStream_t stream) {
    const size_t cbase = p.shared_coef ? 0 : base;
    BatchLoc<CT> loc;
    loc.a = static_cast<const CT*>(p.a) + cbase;
    loc.b = HAS_B ? static_cast<const CT*>(p.b) + cbase : nullptr;
    loc.c = static_cast<const CT*>(p.c) + cbase;
    loc.d = p.d + base;
    loc.x = p.x + base;
    loc.rows = p.rows;
    loc.xed = p.xed;
    loc.n = p.n;
    loc.m = m;
    loc.cstride = p.shared_coef ? 0 : (size_t)p.n;
    hipLaunchKernelGGL((rows_kernel<E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>), dim3(grid), dim3(threads), 0, stream,
                       loc, lg_m, nslices);
    // K2 touches only the rows and x_ed: one instantiation for every storage type
    BatchLoc<float> rl{};
    rl.rows = p.rows;
    rl.xed = p.xed;
    rl.n = p.n;
    rl.m = m;
    hipLaunchKernelGGL((reduced_kernel<METHOD, BatchLoc<float>>), dim3(cs), dim3(rthreads), rshmem, stream, rl);
    hipLaunchKernelGGL((stage3_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>), dim3(grid), dim3(threads),
                       0, stream, loc, lg_m, nslices);
}

template <int METHOD, int E, int L, int MODE>
hipError_t launch_split_el(const SplitArgs& p, hipStream_t stream) {
    const int m = p.n / (E * L);
    const int lg_m = __builtin_ctz((unsigned)m);
    const int threads = 256;
    const int slices_per_block = threads / L;
    static std::atomic<unsigned long long> lds_set{0};  // K2 keeps 10 rows of m floats in LDS
    hipError_t le = ensure_lds_limit(reinterpret_cast<const void*>(reduced_kernel<METHOD, BatchLoc<float>>),
                                     (int)(sizeof(float) * 10 * kSplitMaxM), lds_set);
    if (le != hipSuccess) return le;
    int rthreads = m < 1024 ? m : 1024;
    rthreads = (rthreads + 63) / 64 * 64;
    const size_t rshmem = sizeof(float) * 10 * (size_t)m;
    for (int s0 = 0; s0 < p.nsys; s0 += p.chunk) {
        const int cs = (p.nsys - s0 < p.chunk) ? p.nsys - s0 : p.chunk;
        const size_t base = (size_t)s0 * p.n;
        const long long nslices = (long long)cs * m;
        const int grid = (int)((nslices + slices_per_block - 1) / slices_per_block);
        const bool hb = p.b != nullptr;
        if (p.f16) {
            if (hb) launch_round<METHOD, E, L, MODE, true, _Float16, false>(p, base, grid, threads, lg_m, nslices, cs, m, rthreads, rshmem, stream);
            else launch_round<METHOD, E, L, MODE, false, _Float16, false>(p, base, grid, threads, lg_m, nslices, cs, m, rthreads, rshmem, stream);
        } else if (p.refine) {
            if (hb) launch_round<METHOD, E, L, MODE, true, float, true>(p, base, grid, threads, lg_m, nslices, cs, m, rthreads, rshmem, stream);
            else launch_round<METHOD, E, L, MODE, false, float, true>(p, base, grid, threads, lg_m, nslices, cs, m, rthreads, rshmem, stream);
        } else {
            if (hb) launch_round<METHOD, E, L, MODE, true, float, false>(p, base, grid, threads, lg_m, nslices, cs, m, rthreads, rshmem, stream);
            else launch_round<METHOD, E, L, MODE, false, float, false>(p, base, grid, threads, lg_m, nslices, cs, m, rthreads, rshmem, stream);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int METHOD, int MODE>
hipError_t launch_split_mode(const SplitArgs& p, int E, int L, hipStream_t stream) {
#define TPRH_SPLIT_EL(EE, LL) \
    if (E == EE && L == LL) return launch_split_el<METHOD, EE, LL, MODE>(p, stream);
    TPRH_SPLIT_SHAPES(TPRH_SPLIT_EL)
#undef TPRH_SPLIT_EL
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_split(int method, int mode, const SplitArgs& p, int E, int L, hipStream_t stream) {
    if (method == 0) {
        return mode == kExact ? launch_split_mode<0, kExact>(p, E, L, stream)
                              : launch_split_mode<0, kFast>(p, E, L, stream);
    }
    return mode == kExact ? launch_split_mode<1, kExact>(p, E, L, stream) : launch_split_mode<1, kFast>(p, E, L, stream);
}

}  // namespace tprh
