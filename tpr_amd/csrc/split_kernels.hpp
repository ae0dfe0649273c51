// split_kernels.hpp -- the three kernels of the multi-workgroup (split)
// path, written against a locator of each system's arrays so that the batch
// layout (split.hip) and the serving loop's per-system table (serve_split.hip)
// share them.  See split.hip for the algorithm.
#pragma once
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

// Where system q of a round keeps its arrays.  The kernels below are written
// against this interface so that one body serves the batch layout (systems
// back to back, BatchLoc) and the serving loop's per-system table (TableLoc,
// launch_serve_split).
template <typename CT>
struct BatchLoc {
    const CT *a, *b, *c;  // b may be null (unit diagonal)
    const float* d;
    float* x;
    float* rows;
    float* xed;
    int n, m;
    size_t cstride;  // coefficient stride between systems: n, or 0 for one shared matrix
    __device__ const CT* A(long long q) const { return a + (size_t)q * cstride; }
    __device__ const CT* B(long long q) const { return b ? b + (size_t)q * cstride : nullptr; }
    __device__ const CT* C(long long q) const { return c + (size_t)q * cstride; }
    __device__ const float* D(long long q) const { return d + (size_t)q * n; }
    __device__ float* X(long long q) const { return x + (size_t)q * n; }
    __device__ float* Rows(long long q) const { return rows + (size_t)q * m * kRowWords; }
    __device__ float* Xed(long long q) const { return xed + (size_t)q * m; }
};

// Slice gs of the round: system q, slice j inside it, and the lane's first
// equation lo inside the system.
template <int E, int L>
__device__ __forceinline__ void slice_index(long long gs, int lg_m, int l, long long& q, int& j, size_t& lo) {
    q = gs >> lg_m;
    j = (int)(gs & ((1LL << lg_m) - 1));
    lo = (size_t)j * (E * L) + (size_t)l * E;
}

template <int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE, typename Loc>
__global__ __launch_bounds__(256) void rows_kernel(Loc loc, int lg_m, long long nslices) {
    const int l = threadIdx.x % L;
    const long long g = (long long)blockIdx.x * (blockDim.x / L) + threadIdx.x / L;  // slice over the round
    const bool valid = g < nslices;
    long long q;
    int j;
    size_t lo;
    slice_index<E, L>(valid ? g : 0, lg_m, l, q, j, lo);
    RawPart<E> raw;
    load_raw_t<E, CT>(loc.A(q), HAS_B ? loc.B(q) : nullptr, loc.C(q), loc.D(q), lo, lo, raw);
    float* row = loc.Rows(q) + (size_t)j * kRowWords;
    if constexpr (REFINE) {
        float xo[E], xl, xr;
        load_x_with_halo<E>(loc.X(q), 0, lo, loc.n, xo, xl, xr);
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

template <int METHOD, typename Loc>
__global__ __launch_bounds__(1024) void reduced_kernel(Loc loc) {
    extern __shared__ float lds[];
    const int m = loc.m;
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
    const float* sysrows = loc.Rows(blockIdx.x);
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
    float* xed = loc.Xed(blockIdx.x);
    for (int j = threadIdx.x; j < m; j += nthr) xed[j] = res[j];
}

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE, typename Loc>
__global__ __launch_bounds__(256) void stage3_kernel(Loc loc, int lg_m, long long nslices) {
    const int l = threadIdx.x % L;
    // workgroups take the slices in the reverse of K1's order: the slices K1
    // read last are the likeliest still in the device's last-level cache
    // when a batch is larger than that cache
    const long long g = (long long)(gridDim.x - 1 - blockIdx.x) * (blockDim.x / L) + threadIdx.x / L;
    const bool valid = g < nslices;
    long long q;
    int j;
    size_t lo;
    slice_index<E, L>(valid ? g : 0, lg_m, l, q, j, lo);
    RawPart<E> raw;
    load_raw_t<E, CT>(loc.A(q), HAS_B ? loc.B(q) : nullptr, loc.C(q), loc.D(q), lo, lo, raw);
    float* x = loc.X(q);
    const float* rows = loc.Rows(q);
    float xo[E];
    if constexpr (REFINE) {
        // neighbours inside the slice belong to this wave, whose loads all
        // precede its stores; across slice edges, the values K1 saved
        load_e<E>(x + lo, xo);
        const int m = 1 << lg_m;
        float xl, xr;
        if (l > 0) xl = x[lo - 1];
        else xl = (j > 0) ? rows[(size_t)(j - 1) * kRowWords + 7] : 0.0f;
        if (l < L - 1) xr = x[lo + E];
        else xr = (j < m - 1) ? rows[(size_t)(j + 1) * kRowWords + 6] : 0.0f;
        residual_fp64<E, HAS_B>(xo, xl, xr, raw);
    }
    float A[E], C[E], R[E], X[E];
    normalise<E, MODE>(raw, HAS_B, A, C, R);
    const float* xed = loc.Xed(q);
    const float xe = xed[j];
    const float xprev = (j > 0) ? xed[j - 1] : 0.0f;
    if constexpr (METHOD == 0) {
        ptpr_stage1<E, L, MODE>(A, C, R, l);
        const float* row = rows + (size_t)j * kRowWords;
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
    if (valid) store_e<E>(x + lo, X);
}

}  // namespace tprh
