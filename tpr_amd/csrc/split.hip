// split.hip -- PTPR / TPR for systems larger than one workgroup.
//
// The reference needs every block of a system co-resident for its grid-wide
// barriers and limits n to s^2 (src/ptpr.cu:543,581,655).  Here a system of m
// slices is solved in three kernels, each an ordinary launch:
//   K1 stage1_kernel : every lane group runs stage 1 on its slices
//                      (src/ptpr.cpp:130-190 / src/tpr.cpp:123-216) and
//                      writes the per-equation stage-3 coefficients to a
//                      scratch array plus two boundary rows per slice (E_st
//                      before its update, E_ed);
//   K2 reduced_kernel: one workgroup per system applies E_st += s1 E_ed
//                      (src/ptpr.cpp:193-211), builds the m reduced rows
//                      (:218-241) and solves them in LDS (PCR for PTPR,
//                      src/pcr.cpp; CR for TPR, src/cr.cpp);
//   K3 stage3_kernel : every lane group reads back its coefficients and the
//                      slice-end unknowns and runs stage 3
//                      (src/ptpr.cpp:263-287 / src/tpr.cpp:277-308).
// The batch is processed in rounds of `chunk` systems that reuse one scratch
// area, sized so that a round's inputs and scratch stay in the 256 MB
// Infinity Cache between K1 and K3.
#include <hip/hip_runtime.h>

#include "launch_split.hpp"
#include "slice_solvers.hpp"

namespace tprh {

// Row layout of the boundary buffer: per slice 6 floats,
// [E_st.a, E_st.c, E_st.r, E_ed.a, E_ed.c, E_ed.r].
constexpr int kRowWords = 6;

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
__global__ __launch_bounds__(256) void stage1_kernel(const CT* __restrict__ a, const CT* __restrict__ b,
                                                     const CT* __restrict__ c, const float* __restrict__ d,
                                                     const float* __restrict__ xin, float* __restrict__ sA,
                                                     float* __restrict__ sC, float* __restrict__ sR,
                                                     float* __restrict__ rows, int n, int lg_m, long long nslices) {
    constexpr int S = E * L;
    const int l = threadIdx.x % L;
    const long long g = (long long)blockIdx.x * (blockDim.x / L) + threadIdx.x / L;  // slice over the batch
    const bool valid = g < nslices;
    const long long gs = valid ? g : 0;
    const long long sys = gs >> lg_m;
    const size_t x0 = (size_t)sys * (size_t)n;
    const size_t off = (size_t)gs * S + (size_t)l * E;
    RawPart<E> raw;
    load_raw_t<E, CT>(a, HAS_B ? b : nullptr, c, d, off, raw);
    if constexpr (REFINE) {
        float xo[E];
        residual_fp64<E, HAS_B>(xin, x0, off, n, raw, xo);
    }
    float A[E], C[E], R[E];
    normalise<E, MODE>(raw, HAS_B, A, C, R);
    if constexpr (METHOD == 0) {
        ptpr_stage1<E, L, MODE>(A, C, R, l);
        if (valid) {
            store_e<E>(sA + off, A);
            store_e<E>(sC + off, C);
            store_e<E>(sR + off, R);
        }
    } else {
        float Ao[E], Co[E], Ro[E];
#pragma unroll
        for (int i = 0; i < E; i++) {
            Ao[i] = A[i];
            Co[i] = C[i];
            Ro[i] = R[i];
        }
        tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
        // stage-3 coefficients: originals at even offsets, stage-1 values at
        // odd offsets (src/tpr.cpp:128-132,209-213,381-387)
        float QA[E], QC[E], QR[E];
#pragma unroll
        for (int i = 0; i < E; i++) {
            const bool even = (E % 2 == 0) ? ((i & 1) == 0) : ((l & 1) == 0);
            QA[i] = even ? Ao[i] : A[i];
            QC[i] = even ? Co[i] : C[i];
            QR[i] = even ? Ro[i] : R[i];
        }
        if (valid) {
            store_e<E>(sA + off, QA);
            store_e<E>(sC + off, QC);
            store_e<E>(sR + off, QR);
        }
    }
    if (valid) {
        float* row = rows + (size_t)gs * kRowWords;
        if (l == 0) {
            row[0] = A[0];
            row[1] = C[0];
            row[2] = R[0];
        }
        if (l == L - 1) {
            row[3] = A[E - 1];
            row[4] = C[E - 1];
            row[5] = R[E - 1];
        }
    }
}

template <int METHOD>
__global__ __launch_bounds__(1024) void reduced_kernel(const float* __restrict__ rows, float* __restrict__ xed,
                                                       int m) {
    extern __shared__ float lds[];
    float* st_a = lds;
    float* st_c = st_a + m;
    float* st_r = st_c + m;
    float* ra0 = st_r + m;
    float* rc0 = ra0 + m;
    float* rr0 = rc0 + m;
    float* ra1 = rr0 + m;
    float* rc1 = ra1 + m;
    float* rr1 = rc1 + m;
    float* xo = rr1 + m;
    const float* sysrows = rows + (size_t)blockIdx.x * m * kRowWords;
    const int nthr = blockDim.x;
    for (int j = threadIdx.x; j < m; j += nthr) {
        // E_st += s1 * E_ed, src/ptpr.cpp:193-211 (src/tpr.cpp:327-340)
        const float* r = sysrows + (size_t)j * kRowWords;
        const float c_ed = r[4];
        const float s1 = (fabsf(c_ed) < FLT_EPSILON) ? -1.0f : -r[1] / c_ed;
        st_a[j] = r[0] + s1 * r[3];
        st_c[j] = s1;
        st_r[j] = r[2] + s1 * r[5];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += nthr) {
        // reduced row j from (E_ed(j), E_st(j+1)), src/ptpr.cpp:218-241
        const float* r = sysrows + (size_t)j * kRowWords;
        const float ak = r[3], ck = r[4], rk = r[5];
        if (j < m - 1) {
            const float akr = st_a[j + 1], ckr = st_c[j + 1], rkr = st_r[j + 1];
            const float inv = 1.0f / (1.0f - akr * ck);
            ra0[j] = inv * ak;
            rc0[j] = -inv * ckr * ck;
            rr0[j] = inv * (rk - rkr * ck);
        } else {
            ra0[j] = ak;
            rc0[j] = ck;
            rr0[j] = rk;
        }
    }
    __syncthreads();
    const float* res;
    if constexpr (METHOD == 0) {
        res = lds_pcr(ra0, rc0, rr0, ra1, rc1, rr1, m, threadIdx.x, nthr);
    } else {
        lds_cr(ra0, rc0, rr0, xo, m, threadIdx.x, nthr);
        res = xo;
    }
    for (int j = threadIdx.x; j < m; j += nthr) xed[(size_t)blockIdx.x * m + j] = res[j];
}

template <int METHOD, int E, int L, int MODE, bool REFINE>
__global__ __launch_bounds__(256) void stage3_kernel(const float* __restrict__ sA, const float* __restrict__ sC,
                                                     const float* __restrict__ sR, const float* __restrict__ rows,
                                                     const float* __restrict__ xed, float* __restrict__ x, int lg_m,
                                                     long long nslices) {
    constexpr int S = E * L;
    const int l = threadIdx.x % L;
    const long long g = (long long)blockIdx.x * (blockDim.x / L) + threadIdx.x / L;
    const bool valid = g < nslices;
    const long long gs = valid ? g : 0;
    const int j = (int)(gs & ((1LL << lg_m) - 1));  // slice inside its system
    const size_t off = (size_t)gs * S + (size_t)l * E;
    float A[E], C[E], R[E], X[E];
    load_e<E>(sA + off, A);
    load_e<E>(sC + off, C);
    load_e<E>(sR + off, R);
    const float xe = xed[gs];
    const float xprev = (j > 0) ? xed[gs - 1] : 0.0f;
    if constexpr (METHOD == 0) {
        const float* row = rows + (size_t)gs * kRowWords;
        ptpr_stage3<E, L, MODE>(A, C, R, xprev, xe, SliceEnd{row[3], row[4], row[5]}, l, X);
    } else {
        tpr_stage3<E, L, MODE>(A, C, R, A, C, R, xprev, xe, l, X);
    }
    if (valid) {
        if constexpr (REFINE) {
            float xo[E];
            load_e<E>(x + off, xo);
#pragma unroll
            for (int i = 0; i < E; i++) X[i] = xo[i] + X[i];
        }
        store_e<E>(x + off, X);
    }
}

namespace {

template <int METHOD, int E, int L, int MODE>
hipError_t launch_split_el(const SplitArgs& p, hipStream_t stream) {
    const int m = p.n / (E * L);
    const int lg_m = __builtin_ctz((unsigned)m);
    const int threads = 256;
    const int slices_per_block = threads / L;
    static bool attr_done = false;
    if (!attr_done) {
        const int bytes = (int)(sizeof(float) * 10 * kSplitMaxM);
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(reduced_kernel<METHOD>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) return e;
        attr_done = true;
    }
    int rthreads = m < 1024 ? m : 1024;
    rthreads = (rthreads + 63) / 64 * 64;
    const size_t rshmem = sizeof(float) * 10 * (size_t)m;
    for (int s0 = 0; s0 < p.nsys; s0 += p.chunk) {
        const int cs = (p.nsys - s0 < p.chunk) ? p.nsys - s0 : p.chunk;
        const size_t base = (size_t)s0 * p.n;
        const long long nslices = (long long)cs * m;
        const int grid = (int)((nslices + slices_per_block - 1) / slices_per_block);
        float* x = p.x + base;
        const float* d = p.d + base;
        if (p.f16) {
            const _Float16* a = static_cast<const _Float16*>(p.a) + base;
            const _Float16* c = static_cast<const _Float16*>(p.c) + base;
            const _Float16* b = p.b ? static_cast<const _Float16*>(p.b) + base : nullptr;
            if (b) {
                hipLaunchKernelGGL((stage1_kernel<METHOD, E, L, MODE, true, _Float16, false>), dim3(grid),
                                   dim3(threads), 0, stream, a, b, c, d, nullptr, p.sA, p.sC, p.sR, p.rows, p.n, lg_m,
                                   nslices);
            } else {
                hipLaunchKernelGGL((stage1_kernel<METHOD, E, L, MODE, false, _Float16, false>), dim3(grid),
                                   dim3(threads), 0, stream, a, b, c, d, nullptr, p.sA, p.sC, p.sR, p.rows, p.n,
                                   lg_m, nslices);
            }
        } else {
            const float* a = static_cast<const float*>(p.a) + base;
            const float* c = static_cast<const float*>(p.c) + base;
            const float* b = p.b ? static_cast<const float*>(p.b) + base : nullptr;
            if (p.refine) {
                if (b) {
                    hipLaunchKernelGGL((stage1_kernel<METHOD, E, L, MODE, true, float, true>), dim3(grid),
                                       dim3(threads), 0, stream, a, b, c, d, x, p.sA, p.sC, p.sR, p.rows, p.n, lg_m,
                                       nslices);
                } else {
                    hipLaunchKernelGGL((stage1_kernel<METHOD, E, L, MODE, false, float, true>), dim3(grid),
                                       dim3(threads), 0, stream, a, b, c, d, x, p.sA, p.sC, p.sR, p.rows, p.n, lg_m,
                                       nslices);
                }
            } else if (b) {
                hipLaunchKernelGGL((stage1_kernel<METHOD, E, L, MODE, true, float, false>), dim3(grid), dim3(threads),
                                   0, stream, a, b, c, d, nullptr, p.sA, p.sC, p.sR, p.rows, p.n, lg_m, nslices);
            } else {
                hipLaunchKernelGGL((stage1_kernel<METHOD, E, L, MODE, false, float, false>), dim3(grid),
                                   dim3(threads), 0, stream, a, b, c, d, nullptr, p.sA, p.sC, p.sR, p.rows, p.n, lg_m,
                                   nslices);
            }
        }
        hipLaunchKernelGGL((reduced_kernel<METHOD>), dim3(cs), dim3(rthreads), rshmem, stream, p.rows, p.xed, m);
        if (p.refine) {
            hipLaunchKernelGGL((stage3_kernel<METHOD, E, L, MODE, true>), dim3(grid), dim3(threads), 0, stream, p.sA,
                               p.sC, p.sR, p.rows, p.xed, x, lg_m, nslices);
        } else {
            hipLaunchKernelGGL((stage3_kernel<METHOD, E, L, MODE, false>), dim3(grid), dim3(threads), 0, stream, p.sA,
                               p.sC, p.sR, p.rows, p.xed, x, lg_m, nslices);
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
