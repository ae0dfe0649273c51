// batched_impl.hpp -- batched PTPR / TPR: many independent systems of the
// same size, one group of lanes per system, one launch for the whole batch.
// Included by batched_ptpr.hip and batched_tpr.hip, one method each, so the
// two halves compile in parallel.
//
// This replaces the reference's one-system-per-launch cooperative kernels
// (src/ptpr.cu:48-150, src/tpr.cu:44-156): a system lives inside one
// workgroup, so there is no grid-wide barrier and no co-residency limit, and
// the reduced system is solved by the system's own workgroup.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <type_traits>

#include "launch.hpp"
#include "slice_solvers.hpp"
#include "staged_io.hpp"

namespace tprh {

// One workgroup holds sys_per_block systems side by side (or one system of up
// to BLOCK lanes).  HAS_B: the diagonal is given and divided out in registers
// (false: unit diagonal, b is not read).  CT: storage type of a, b, c (float
// or _Float16).  REFINE: one iterative-refinement step, x += A^-1 (d - A x)
// with the residual accumulated in FP64.  BLOCK == kSmallBlock: every system
// fits in one wavefront (m * L <= 64) and its reduced solve synchronises
// per wavefront, so the four wavefronts of a workgroup run independently.
constexpr int kSmallBlock = 256;

template <int METHOD, int E, int L, int MODE, int BLOCK, bool HAS_B, typename CT, bool REFINE>
__global__ __launch_bounds__(BLOCK) void batched_kernel(const CT* __restrict__ a, const CT* __restrict__ b,
                                                       const CT* __restrict__ c, const float* __restrict__ d,
                                                       float* __restrict__ x, int n, int m, int nsys,
                                                       int lg_lanes, int shared, int x_cached) {
    extern __shared__ float lds[];
    const int sl = threadIdx.x >> lg_lanes;             // system slot in the workgroup
    const int t = threadIdx.x & ((1 << lg_lanes) - 1);  // lane inside the system
    float* my = lds + sl * lds_words_per_system(m);
    const int sys = blockIdx.x * (blockDim.x >> lg_lanes) + sl;
    const bool valid = sys < nsys;
    const size_t x0 = valid ? (size_t)sys * (size_t)n : 0;
    const size_t off = x0 + (size_t)t * E;
    const size_t coff = shared ? (size_t)t * E : off;  // shared: one matrix for every right-hand side
    // the staged layout: this wavefront's run of 64 E floats and its LDS
    constexpr bool STAGE = kStageX<E>;
    float* stage = lds + stage_base(blockDim.x >> lg_lanes, m) + (threadIdx.x >> 6) * stage_wave_words<E>();
    const size_t wave_off = ((size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) * E;
    const size_t total = (size_t)nsys * (size_t)n;
    RawPart<E> raw;
    if constexpr (STAGE && std::is_same<CT, float>::value) {
        if (!shared) load_raw_staged<E, HAS_B>(a, b, c, d, wave_off, total, stage, threadIdx.x & 63, raw);
        else load_raw_t<E, CT>(a, HAS_B ? b : nullptr, c, d, coff, off, raw);
    } else {
        load_raw_t<E, CT>(a, HAS_B ? b : nullptr, c, d, coff, off, raw);
    }
    float xo[E];
    if constexpr (REFINE) {
        // every lane of the system loads x here, before the barriers of the
        // reduced solve; all stores of x come after them
        float xl, xr;
        load_x_with_halo<E>(x, x0, off, n, xo, xl, xr);
        residual_fp64<E, HAS_B>(xo, xl, xr, raw);
    }
    float X[E];
    constexpr bool WAVE = BLOCK == kSmallBlock;
    if constexpr (METHOD == 0) {
        ptpr_solve_raw<E, L, MODE, WAVE>(raw, HAS_B, m, t, my, X);
    } else {
        float A[E], C[E], R[E];
        normalise<E, MODE>(raw, HAS_B, A, C, R);
        tpr_solve_regs<E, L, MODE, WAVE>(A, C, R, m, t, my, X);
    }
    if constexpr (REFINE) {
#pragma unroll
        for (int i = 0; i < E; i++) X[i] = xo[i] + X[i];
    }
    // x is not read again by this launch: non-temporal stores by default
    // (4096 x N=1024 with staged stores: 16.5 against 17.2 us on one stream,
    // 13.5 against 13.6 us per batch on three); x_cached keeps x in the
    // caches for a caller whose next kernel reads it
    if constexpr (STAGE) {
        // every lane takes part (the pieces past the batch are not stored)
        if (x_cached) store_x_staged<E, false>(x, X, stage, wave_off, total, threadIdx.x & 63);
        else store_x_staged<E, true>(x, X, stage, wave_off, total, threadIdx.x & 63);
    } else if (valid) {
        if (x_cached) store_e<E>(x + off, X);
        else store_e<E, true>(x + off, X);
    }
}

namespace {

template <int METHOD, int E, int L, int MODE, int BLOCK, typename CT, bool REFINE>
hipError_t launch_b(const BatchArgs& p, int grid, int block, size_t shmem, int lg, hipStream_t stream) {
    const CT* a = static_cast<const CT*>(p.a);
    const CT* b = static_cast<const CT*>(p.b);
    const CT* c = static_cast<const CT*>(p.c);
    auto* kern = p.b != nullptr ? batched_kernel<METHOD, E, L, MODE, BLOCK, true, CT, REFINE>
                                : batched_kernel<METHOD, E, L, MODE, BLOCK, false, CT, REFINE>;
    if (shmem > 65536) {
        // 1024-lane systems of 16 equations per lane stage 68 KB
        static std::atomic<unsigned long long> set_b{0}, set_u{0};
        hipError_t e = ensure_lds_limit(reinterpret_cast<const void*>(kern), (int)shmem, p.b != nullptr ? set_b : set_u);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shmem, stream, a, b, c, p.d, p.x, p.n, p.m, p.nsys, lg,
                       p.shared_coef ? 1 : 0, p.x_cached ? 1 : 0);
    return hipSuccess;
}

template <int METHOD, int E, int L, int MODE, int BLOCK>
hipError_t launch_variant(const BatchArgs& p, int grid, int block, size_t shmem, int lg, hipStream_t stream) {
    if (!p.f16 && !p.refine) {
        return launch_b<METHOD, E, L, MODE, BLOCK, float, false>(p, grid, block, shmem, lg, stream);
    }
    // the storage and refinement variants exist for the default shapes only
    if constexpr (is_default_shape(E, L)) {
        if (p.f16 && !p.refine) {
            return launch_b<METHOD, E, L, MODE, BLOCK, _Float16, false>(p, grid, block, shmem, lg, stream);
        }
        if (!p.f16 && p.refine) {
            return launch_b<METHOD, E, L, MODE, BLOCK, float, true>(p, grid, block, shmem, lg, stream);
        }
    }
    return hipErrorInvalidValue;
}

template <int METHOD, int E, int L, int MODE>
hipError_t launch_el(const BatchArgs& p, hipStream_t stream) {
    const int T = p.m * L;  // lanes per system, a power of two
    const int lg = __builtin_ctz((unsigned)T);
    const int block = T <= 64 ? kSmallBlock : T;
    const int spb = block / T;
    const int grid = (p.nsys + spb - 1) / spb;
    size_t shmem = sizeof(float) * (size_t)spb * lds_words_per_system(p.m);
    if (kStageX<E>)
        shmem = sizeof(float) * ((size_t)stage_base(spb, p.m) + (size_t)stage_words<E>(block));
    // the wavefront-synchronised variant only when a system fits one
    // wavefront; systems of up to 512 lanes take a kernel bounded to 512
    // lanes per workgroup, whose waves may hold up to 256 VGPRs (a bound of
    // 1024 lanes caps them at 128, and exact arithmetic on 16 equations per
    // lane then spills)
    hipError_t e = T <= 64    ? launch_variant<METHOD, E, L, MODE, kSmallBlock>(p, grid, block, shmem, lg, stream)
                   : T <= 512 ? launch_variant<METHOD, E, L, MODE, 512>(p, grid, block, shmem, lg, stream)
                              : launch_variant<METHOD, E, L, MODE, 1024>(p, grid, block, shmem, lg, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

template <int METHOD, int MODE>
hipError_t launch_mode(const BatchArgs& p, int E, int L, hipStream_t stream) {
#define TPRH_EL(EE, LL) \
    if (E == EE && L == LL) return launch_el<METHOD, EE, LL, MODE>(p, stream);
    TPRH_SLICE_SHAPES(TPRH_EL)
#undef TPRH_EL
    return hipErrorInvalidValue;
}

}  // namespace

template <int METHOD>
hipError_t launch_batched_method(int mode, const BatchArgs& p, int E, int L, hipStream_t stream) {
    return mode == kExact ? launch_mode<METHOD, kExact>(p, E, L, stream) : launch_mode<METHOD, kFast>(p, E, L, stream);
}

}  // namespace tprh
