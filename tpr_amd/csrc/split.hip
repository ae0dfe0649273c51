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

#include "launch_split.hpp"
#include "slice_solvers.hpp"

namespace tprh {

// Row layout of the boundary buffer: per slice 8 floats,
// [E_st.a, E_st.c, E_st.r, E_ed.a, E_ed.c, E_ed.r, x_st, x_ed]; the last two
// (refinement only) keep the slice's first and last x of the current iterate,
// so that K3 can rebuild the residual at slice edges after neighbouring
// workgroups have overwritten x.
constexpr int kRowWords = kSplitRowWords;

template <int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
__global__ __launch_bounds__(256) void rows_kernel(const CT* __restrict__ a, const CT* __restrict__ b,
                                                   const CT* __restrict__ c, const float* __restrict__ d,
                                                   const float* __restrict__ xin, float* __restrict__ rows, int n,
                                                   int lg_m, long long nslices, size_t cstride) {
    constexpr int S = E * L;
    const int l = threadIdx.x % L;
    const long long g = (long long)blockIdx.x * (blockDim.x / L) + threadIdx.x / L;  // slice over the batch
    const bool valid = g < nslices;
    const long long gs = valid ? g : 0;
    const size_t x0 = (size_t)(gs >> lg_m) * (size_t)n;
    const size_t off = (size_t)gs * S + (size_t)l * E;
    const size_t coff = (size_t)(gs >> lg_m) * cstride + (off - x0);  // cstride 0: one shared matrix
    RawPart<E> raw;
    load_raw_t<E, CT>(a, HAS_B ? b : nullptr, c, d, coff, off, raw);
    float* row = rows + (size_t)gs * kRowWords;
    if constexpr (REFINE) {
        float xo[E], xl, xr;
        load_x_with_halo<E>(xin, x0, off, n, xo, xl, xr);
        residual_fp64<E, HAS_B>(xo, xl, xr, raw);
        if (valid && l == 0) row[6] = xo[0];
        if (valid && l == L - 1) row[7] = xo[E - 1];
    }
    float A[E], C[E], R[E];
    normalise<E, MODE>(raw, HAS_B, A, C, R);
    tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
    if (valid) {
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

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
__global__ __launch_bounds__(256) void stage3_kernel(const CT* __restrict__ a, const CT* __restrict__ b,
                                                     const CT* __restrict__ c, const float* __restrict__ d,
                                                     const float* __restrict__ rows, const float* __restrict__ xed,
                                                     float* __restrict__ x, int n, int lg_m, long long nslices,
                                                     size_t cstride) {
    constexpr int S = E * L;
    const int l = threadIdx.x % L;
    const long long g = (long long)blockIdx.x * (blockDim.x / L) + threadIdx.x / L;
    const bool valid = g < nslices;
    const long long gs = valid ? g : 0;
    const int j = (int)(gs & ((1LL << lg_m) - 1));  // slice inside its system
    const size_t x0 = (size_t)(gs >> lg_m) * (size_t)n;
    const size_t off = (size_t)gs * S + (size_t)l * E;
    const size_t coff = (size_t)(gs >> lg_m) * cstride + (off - x0);  // cstride 0: one shared matrix
    RawPart<E> raw;
    load_raw_t<E, CT>(a, HAS_B ? b : nullptr, c, d, coff, off, raw);
    float xo[E];
    if constexpr (REFINE) {
        // neighbours inside the slice belong to this wave, whose loads all
        // precede its stores; across slice edges, the values K1 saved
        load_e<E>(x + off, xo);
        const int m = 1 << lg_m;
        float xl, xr;
        if (l > 0) xl = x[off - 1];
        else xl = (j > 0) ? rows[(size_t)(gs - 1) * kRowWords + 7] : 0.0f;
        if (l < L - 1) xr = x[off + E];
        else xr = (j < m - 1) ? rows[(size_t)(gs + 1) * kRowWords + 6] : 0.0f;
        residual_fp64<E, HAS_B>(xo, xl, xr, raw);
    }
    float A[E], C[E], R[E], X[E];
    normalise<E, MODE>(raw, HAS_B, A, C, R);
    const float xe = xed[gs];
    const float xprev = (j > 0) ? xed[gs - 1] : 0.0f;
    if constexpr (METHOD == 0) {
        ptpr_stage1<E, L, MODE>(A, C, R, l);
        const float* row = rows + (size_t)gs * kRowWords;
        ptpr_stage3<E, L, MODE>(A, C, R, xprev, xe, SliceEnd{row[3], row[4], row[5]}, l, X);
    } else {
        float Ao[E], Co[E], Ro[E];
#pragma unroll
        for (int i = 0; i < E; i++) {
            Ao[i] = A[i];
            Co[i] = C[i];
            Ro[i] = R[i];
        }
        tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
        tpr_stage3<E, L, MODE>(A, C, R, Ao, Co, Ro, xprev, xe, l, X);
    }
    if constexpr (REFINE) {
#pragma unroll
        for (int i = 0; i < E; i++) X[i] = xo[i] + X[i];
    }
    if (valid) store_e<E>(x + off, X);
}

namespace {

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
void launch_round(const SplitArgs& p, size_t base, int grid, int threads, int lg_m, long long nslices, int cs, int m,
                  int rthreads, size_t rshmem, hipStream_t stream) {
    const size_t cbase = p.shared_coef ? 0 : base;
    const size_t cstride = p.shared_coef ? 0 : (size_t)p.n;
    const CT* a = static_cast<const CT*>(p.a) + cbase;
    const CT* c = static_cast<const CT*>(p.c) + cbase;
    const CT* b = HAS_B ? static_cast<const CT*>(p.b) + cbase : nullptr;
    const float* d = p.d + base;
    float* x = p.x + base;
    hipLaunchKernelGGL((rows_kernel<E, L, MODE, HAS_B, CT, REFINE>), dim3(grid), dim3(threads), 0, stream, a, b, c, d,
                       x, p.rows, p.n, lg_m, nslices, cstride);
    hipLaunchKernelGGL((reduced_kernel<METHOD>), dim3(cs), dim3(rthreads), rshmem, stream, p.rows, p.xed, m);
    hipLaunchKernelGGL((stage3_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE>), dim3(grid), dim3(threads), 0, stream,
                       a, b, c, d, p.rows, p.xed, x, p.n, lg_m, nslices, cstride);
}

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
