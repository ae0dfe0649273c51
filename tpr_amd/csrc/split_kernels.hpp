// split_kernels.hpp -- the three kernels of the multi-workgroup (split)
// path, written against a locator of each system's arrays so that the batch
// layout (split.hip) and the serving loop's per-system table (serve_split.hip)
// share them.  See split.hip for the algorithm.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "launch_split.hpp"
#include "slice_solvers.hpp"
#include "staged_io.hpp"

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

// Stage 1 in fast arithmetic for PTPR on 16-lane slices takes the two-level
// form (a Thomas sweep in each lane, PCR across the lanes' last rows), as the
// one-workgroup path does; everything else the reference's order.
template <int METHOD, int E, int L, int MODE>
constexpr bool kSplitTwoLevel = METHOD == 0 && MODE == kFast && L == 16 && E >= 2;

// K1 and K3 load (and K3 stores) their slices as 1 KB runs per instruction
// through LDS (staged_io.hpp) where the shape stages and the coefficients are
// fp32: a wavefront holds 64 / L whole slices of one system (m L > 1024), one
// contiguous run starting at `run`.  256-lane workgroups: four wavefronts.
// K3 stores x non-temporally (256 x N=65536: 0.0951-0.0955 against
// 0.0965-0.0968 ms with cached stores).
template <int E, typename CT>
constexpr bool kSplitStaged = kStageX<E> && std::is_same<CT, float>::value;
template <int E, typename CT>
constexpr int kSplitStageWords = kSplitStaged<E, CT> ? 4 * stage_wave_words<E>() : 4;

template <int E, bool HAS_B, typename CT, typename Loc>
__device__ __forceinline__ void load_slices(const Loc& loc, long long q, size_t lo, size_t run, float* stage,
                                            RawPart<E>& raw) {
    if constexpr (kSplitStaged<E, CT>) {
        load_raw_staged<E, HAS_B>(loc.A(q), HAS_B ? loc.B(q) : nullptr, loc.C(q), loc.D(q), run, (size_t)loc.n,
                                  stage, threadIdx.x & 63, raw);
    } else {
        load_raw_t<E, CT>(loc.A(q), HAS_B ? loc.B(q) : nullptr, loc.C(q), loc.D(q), lo, lo, raw);
    }
}

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE, typename Loc>
__global__ __launch_bounds__(256) void rows_kernel(Loc loc, int lg_m, long long nslices) {
    __shared__ __align__(16) float stage_lds[kSplitStageWords<E, CT>];
    const int l = threadIdx.x % L;
    const long long g = (long long)blockIdx.x * (blockDim.x / L) + threadIdx.x / L;  // slice over the round
    const bool valid = g < nslices;
    long long q;
    int j;
    size_t lo;
    slice_index<E, L>(valid ? g : 0, lg_m, l, q, j, lo);
    // valid is uniform across a wavefront (nslices is a multiple of m)
    const size_t run = valid ? lo - (size_t)(threadIdx.x & 63) * E : 0;
    float* stage = kSplitStaged<E, CT> ? stage_lds + (threadIdx.x >> 6) * stage_wave_words<E>() : stage_lds;
    RawPart<E> raw;
    load_slices<E, HAS_B, CT>(loc, q, lo, run, stage, raw);
    float* row = loc.Rows(q) + (size_t)j * kRowWords;
    if constexpr (REFINE) {
        float xo[E], xl, xr;
        load_x_with_halo<E>(loc.X(q), 0, lo, loc.n, xo, xl, xr);
        residual_fp64<E, HAS_B>(xo, xl, xr, raw);
        if (valid && l == 0) row[6] = xo[0];
        if (valid && l == L - 1) row[7] = xo[E - 1];
    }
    float st_a, st_c, st_r, ed_a, ed_c, ed_r;
    if constexpr (kSplitTwoLevel<METHOD, E, L, MODE>) {
        Fast16State<E> fs;
        fast16_stage1<E>(raw, HAS_B, l, fs);
        st_a = fs.sA, st_c = fs.sC, st_r = fs.sR;
        ed_a = fs.eA, ed_c = fs.eC, ed_r = fs.eR;
    } else {
        float A[E], C[E], R[E];
        normalise<E, MODE>(raw, HAS_B, A, C, R);
        tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
        st_a = A[0], st_c = C[0], st_r = R[0];
        ed_a = A[E - 1], ed_c = C[E - 1], ed_r = R[E - 1];
    }
    if (valid) {
        if (l == 0) {
            row[0] = st_a;
            row[1] = st_c;
            row[2] = st_r;
        }
        if (l == L - 1) {
            row[3] = ed_a;
            row[4] = ed_c;
            row[5] = ed_r;
        }
    }
}

// reduced_rows_lds: the reduced system of one system from its m slices'
// boundary rows (sysrows, kRowWords floats per slice): E_st += s1 E_ed
// (src/ptpr.cpp:193-211, src/tpr.cpp:327-340), the reduced rows
// (src/ptpr.cpp:218-241) and their solve in LDS, PCR for PTPR
// (src/pcr.cpp:15-89) or CR for TPR (src/cr.cpp:23-116), by the nthr threads
// t = 0 .. nthr-1 of one workgroup; returns x at every slice end, in lds.
// lds: 10 m floats.  Every thread of the workgroup must call it.
// Loads and stores of the data blocks exchange inside one launch (boundary
// rows, x at slice ends, counters, flags): atomic at device scope, so that
// they are coherent across the chip's L2 caches without writing back a whole
// L2 (which a device-scope release fence does, and which costs far more than
// these few words).
__device__ __forceinline__ float ld_dev(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int METHOD, bool COHERENT = false>
__device__ __forceinline__ const float* reduced_rows_lds(const float* sysrows, int m, float* lds, int t, int nthr) {
    auto ld = [](const float* p) { return COHERENT ? ld_dev(p) : *p; };
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
    // E_ed waits in ra1 / rc1 / rr1 (free until the solve) between the passes
    for (int j = t; j < m; j += nthr) {
        const float* r = sysrows + (size_t)j * kRowWords;
        const float a_st = ld(r + 0), c_st = ld(r + 1), r_st = ld(r + 2);
        const float a_ed = ld(r + 3), c_ed = ld(r + 4), r_ed = ld(r + 5);
        const float s1 = (fabsf(c_ed) < FLT_EPSILON) ? -1.0f : -c_st / c_ed;
        st_a[j] = a_st + s1 * a_ed;
        st_c[j] = s1;
        st_r[j] = r_st + s1 * r_ed;
        ra1[j] = a_ed;
        rc1[j] = c_ed;
        rr1[j] = r_ed;
    }
    __syncthreads();
    for (int j = t; j < m; j += nthr) {
        // reduced row j from (E_ed(j), E_st(j+1))
        const float ak = ra1[j], ck = rc1[j], rk = rr1[j];
        if (j < m - 1) {
            const float akr = st_a[j + 1], ckr = st_c[j + 1], rkr = st_r[j + 1];
            const float inv = recip_rn(1.0f - akr * ck);
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
    if constexpr (METHOD == 0) {
        return lds_pcr(ra0, rc0, rr0, ra1, rc1, rr1, m, t, nthr);
    } else {
        lds_cr(ra0, rc0, rr0, xo, m, t, nthr);
        return xo;
    }
}

// As reduced_rows_lds, x at every slice end then written to xed.
template <int METHOD, bool COHERENT = false>
__device__ __forceinline__ void reduced_rows_solve(const float* sysrows, float* xed, int m, float* lds, int t,
                                                   int nthr) {
    const float* res = reduced_rows_lds<METHOD, COHERENT>(sysrows, m, lds, t, nthr);
    for (int j = t; j < m; j += nthr) {
        if constexpr (COHERENT) st_dev(xed + j, res[j]);
        else xed[j] = res[j];
    }
}

template <int METHOD, typename Loc>
__global__ __launch_bounds__(1024) void reduced_kernel(Loc loc) {
    extern __shared__ float lds[];
    reduced_rows_solve<METHOD>(loc.Rows(blockIdx.x), loc.Xed(blockIdx.x), loc.m, lds, threadIdx.x, blockDim.x);
}

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE, typename Loc>
__global__ __launch_bounds__(256) void stage3_kernel(Loc loc, int lg_m, long long nslices) {
    __shared__ __align__(16) float stage_lds[kSplitStageWords<E, CT>];
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
    const size_t run = valid ? lo - (size_t)(threadIdx.x & 63) * E : 0;
    float* stage = kSplitStaged<E, CT> ? stage_lds + (threadIdx.x >> 6) * stage_wave_words<E>() : stage_lds;
    RawPart<E> raw;
    load_slices<E, HAS_B, CT>(loc, q, lo, run, stage, raw);
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
    float X[E];
    const float* xed = loc.Xed(q);
    const float xe = xed[j];
    const float xprev = (j > 0) ? xed[j - 1] : 0.0f;
    if constexpr (kSplitTwoLevel<METHOD, E, L, MODE>) {
        Fast16State<E> fs;
        fast16_stage1<E>(raw, HAS_B, l, fs);
        const int src = (int)(threadIdx.x & ~(unsigned)(L - 1)) + (L - 1);  // the slice's last lane
        SliceEnd se;
        se.a = __shfl(fs.eA, src, 64);
        se.c = __shfl(fs.eC, src, 64);
        se.r = __shfl(fs.eR, src, 64);
        fast16_stage3<E>(fs, xprev, xe, se, l, X);
    } else if constexpr (METHOD == 0) {
        float A[E], C[E], R[E];
        normalise<E, MODE>(raw, HAS_B, A, C, R);
        ptpr_stage1<E, L, MODE>(A, C, R, l);
        const float* row = rows + (size_t)j * kRowWords;
        ptpr_stage3<E, L, MODE>(A, C, R, xprev, xe, SliceEnd{row[3], row[4], row[5]}, l, X);
    } else {
        float A[E], C[E], R[E];
        normalise<E, MODE>(raw, HAS_B, A, C, R);
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
    if constexpr (kSplitStaged<E, CT>) {
        if (valid) store_x_staged<E, true>(x, X, stage, run, (size_t)loc.n, threadIdx.x & 63);
    } else if (valid) {
        store_e<E>(x + lo, X);
    }
}

// ---------------------------------------------------------------------------
// One-pass persistent form of the multi-workgroup path: every equation is
// read once and x written once (20 B/eq), against 36 B/eq for the three
// kernels above, whose stage 3 reloads the inputs to redo stage 1.
//
// A block is kPersistThreads lanes = kPersistThreads / L consecutive slices
// of one system; a system of m slices is bps = m L / kPersistThreads blocks.
// Each workgroup of the persistent grid takes tickets t = 0, 1, ... from an
// atomic counter (block t % bps of system t / bps) until they run out, and
// for each block
//   1. loads its slices (vector loads, the division by the diagonal in
//      registers) and runs stage 1 in registers;
//   2. writes each slice's boundary rows (E_st before its update, E_ed) to
//      the workspace and counts itself in the system's arrival counter;
//   3. the last block of the system to arrive solves the reduced system
//      (reduced_rows_solve) and raises the system's flag; the other blocks
//      wait for the flag with their slices still in registers;
//   4. runs stage 3 in registers from x at the slice ends and stores x.
// The wait cannot block forever: tickets are handed out in system order and
// only to workgroups already running, so the oldest unfinished system has
// all of its blocks among the running workgroups as soon as at least bps of
// them are resident (the host launches this form only when the device holds
// that many at once), and its last arriver finishes it.  The wait is also
// bounded (about two seconds) and records a timeout in sync[] if it ever
// expires, so that no workgroup spins for ever.
//
// sync[] (cleared before every launch), one 128-byte line per system so
// that waiting workgroups of different systems poll different lines:
// sync[0] ticket counter, sync[1] wait timeouts, sync[kSyncStride (q + 1)]
// arrivals of system q, sync[kSyncStride (q + 1) + kSyncStride / 2] its flag.
constexpr int kPersistThreads = 256;
constexpr long long kPersistSpinLimit = 1LL << 23;

__device__ __forceinline__ bool wait_flag(int* flag) {
    for (long long i = 0; i < kPersistSpinLimit; i++) {
        if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return true;
        __builtin_amdgcn_s_sleep(8);  // about 0.2 us between polls
    }
    return false;
}

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE, typename Loc>
__global__ __launch_bounds__(kPersistThreads) void persistent_split_kernel(Loc loc, int* __restrict__ sync, int nsys,
                                                                           int lg_bps) {
    extern __shared__ float lds[];  // 10 m floats: the reduced solve
    __shared__ int s_ticket, s_last;
    constexpr int kSlices = kPersistThreads / L;
    const int m = loc.m;
    const int bps = 1 << lg_bps;
    const int total = nsys << lg_bps;
    const int l = threadIdx.x % L;
    const int jb = threadIdx.x / L;  // slice inside the block
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(&sync[0], 1);
        __syncthreads();
        const int tk = s_ticket;
        __syncthreads();  // every thread has read the ticket before it is rewritten
        if (tk >= total) break;  // uniform: the workgroup leaves once the tickets are gone
        const int q = tk >> lg_bps;
        const int j = (tk & (bps - 1)) * kSlices + jb;  // slice inside the system
        const size_t lo = (size_t)j * (E * L) + (size_t)l * E;
        RawPart<E> raw;
        load_raw_t<E, CT>(loc.A(q), HAS_B ? loc.B(q) : nullptr, loc.C(q), loc.D(q), lo, lo, raw);
        float xo[E];
        if constexpr (REFINE) {
            // every block of the system loads its x (and the halo from its
            // neighbours) before the system's last arriver can start the
            // reduced solve, and no x is stored before that solve is done
            float xl, xr;
            load_x_with_halo<E>(loc.X(q), 0, lo, loc.n, xo, xl, xr);
            residual_fp64<E, HAS_B>(xo, xl, xr, raw);
        }
        float* row = loc.Rows(q) + (size_t)j * kRowWords;
        // stage 1, and the slice's boundary rows
        float A[E], C[E], R[E];
        float Ao[E], Co[E], Ro[E];
        Fast16State<(E >= 2 ? E : 2)> fs;
        constexpr bool kTwoLevel = METHOD == 0 && MODE == kFast && L == 16 && E >= 2;
        float st_a, st_c, st_r, ed_a, ed_c, ed_r;
        if constexpr (kTwoLevel) {
            fast16_stage1<E>(raw, HAS_B, l, fs);
            st_a = fs.sA, st_c = fs.sC, st_r = fs.sR;
            ed_a = fs.eA, ed_c = fs.eC, ed_r = fs.eR;
        } else {
            normalise<E, MODE>(raw, HAS_B, A, C, R);
            if constexpr (METHOD == 0) {
                ptpr_stage1<E, L, MODE>(A, C, R, l);
            } else {
#pragma unroll
                for (int i = 0; i < E; i++) {
                    Ao[i] = A[i];
                    Co[i] = C[i];
                    Ro[i] = R[i];
                }
                tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
            }
            st_a = A[0], st_c = C[0], st_r = R[0];
            ed_a = A[E - 1], ed_c = C[E - 1], ed_r = R[E - 1];
        }
        if (l == 0) {
            st_dev(row + 0, st_a);
            st_dev(row + 1, st_c);
            st_dev(row + 2, st_r);
        }
        if (l == L - 1) {
            st_dev(row + 3, ed_a);
            st_dev(row + 4, ed_c);
            st_dev(row + 5, ed_r);
        }
        // publish: every wave's row stores have completed (device-coherent
        // stores, acknowledged once they are visible to the whole chip)
        // before the arrival is counted; the count, the flag and every read
        // of rows and x_ed below are device-coherent too, and each read is
        // issued only after the count or the flag it depends on was seen
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        int* sys_sync = sync + (size_t)kSplitSyncStride * (q + 1);
        if (threadIdx.x == 0)
            s_last = __hip_atomic_fetch_add(sys_sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == bps - 1;
        __syncthreads();
        if (s_last) {
            reduced_rows_solve<METHOD, true>(loc.Rows(q), loc.Xed(q), m, lds, threadIdx.x, kPersistThreads);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(sys_sync + kSplitSyncStride / 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (threadIdx.x == 0 && !wait_flag(sys_sync + kSplitSyncStride / 2))
                __hip_atomic_fetch_add(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
        }
        // stage 3
        const float* xed = loc.Xed(q);
        const float xe = ld_dev(xed + j);
        const float xprev = (j > 0) ? ld_dev(xed + j - 1) : 0.0f;
        float X[E];
        if constexpr (METHOD == 0) {
            // the slice's last row (E_ed) from its last lane
            const int src = (int)(threadIdx.x & ~(unsigned)(L - 1)) + (L - 1);
            SliceEnd se;
            se.a = __shfl(ed_a, src, 64);
            se.c = __shfl(ed_c, src, 64);
            se.r = __shfl(ed_r, src, 64);
            if constexpr (kTwoLevel) {
                fast16_stage3<E>(fs, xprev, xe, se, l, X);
            } else {
                ptpr_stage3<E, L, MODE>(A, C, R, xprev, xe, se, l, X);
            }
        } else {
            tpr_stage3<E, L, MODE>(A, C, R, Ao, Co, Ro, xprev, xe, l, X);
        }
        if constexpr (REFINE) {
#pragma unroll
            for (int i = 0; i < E; i++) X[i] = xo[i] + X[i];
        }
        store_e<E>(loc.X(q) + lo, X);
        // the next ticket's stage 1 may reuse lds only after this block's
        // reduced solve is complete (it is: the flag came after it)
    }
}

// ---------------------------------------------------------------------------
// Pipelined form: one persistent launch per round whose workgroups take
// tasks in ticket order from a counter (sync[0]).  Three kinds of task:
//   rows task    (block i of kPersistThreads / L slices, block i % bps of
//                system i / bps): load the slices, stage 1, write their
//                boundary rows (E_st before its update, E_ed; with
//                refinement also x at the slice edges), count the arrival in
//                the system's counter;
//   reduced task (system q): wait until all bps rows tasks of q have
//                arrived, solve the reduced system (reduced_rows_solve),
//                raise q's flag;
//   stage-3 task (block i): load the slices again, wait for the flag, redo
//                stage 1, run stage 3 and store x.
// Two streams of tasks share the tickets (pipeline_task, launch_split.hpp):
// the rows tasks, and a lag D behind them each system's reduced task
// followed by the stage-3 tasks of the system g before it.  So a system's reduced
// task comes about 2 D tickets after its rows tasks and its stage-3 tasks
// about 2 g (bps + 1) tickets after that: with D about the number of
// workgroups resident at once and g = D / (2 bps), each wait is normally
// over when a task starts, and a stage-3 task's reload finds its slices in
// the device's last-level cache, so HBM sees the inputs about once.
// Device-coherent operations wait behind the memory traffic (microseconds at
// full bandwidth), so no workgroup waits for one where it need not: the next
// ticket is requested as a task starts and read at its end; a stage-3 task
// reads its flag ahead of its loads; a rows task's row stores complete, and
// its arrival is counted, while the workgroup's next task loads (before
// anything in that task waits for another workgroup).
// No wait can block for ever, whatever the residency: every task depends
// only on tasks of lower tickets (rows tasks of q < reduced task of q <
// stage-3 tasks of q, for any D >= 2 bps), a workgroup takes tickets in
// increasing order, holds at most its current and its next one, and counts
// its previous rows task's arrival before it waits for anything, so the
// lowest unfinished ticket is some workgroup's current task and all it
// waits for is done or under way.  The waits are also bounded like the
// one-pass form's, with timeouts counted in sync[1].
// Answers are the same bits as the one-pass form's: the same stage 1 (the
// two-level form for fast PTPR on 16-lane slices; in exact arithmetic the
// O(S) tree for the rows, whose end rows equal the in-slice PCR's bit for
// bit), the same reduced solve and the same stage 3.
__device__ __forceinline__ bool wait_count(int* counter, int target) {
    for (long long i = 0; i < kPersistSpinLimit; i++) {
        if (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        __builtin_amdgcn_s_sleep(8);
    }
    return false;
}

// Points of a pipelined task at which the diagnostic build reads the GPU
// clock (libtpr_diag.so passes a clock that does; this one does nothing and
// compiles away): mark(0) task start, mark(1) loads complete / arrivals
// seen, mark(2) stage 1 done / reduced system solved, mark(3) rows stored /
// flag seen / flag raised, mark(4) x_ed read, mark(5) next ticket read;
// loads_done() makes mark(1) exact, keep() keeps a stage before its mark,
// task() records the ticket's stamps.
struct NoStageClock {
    __device__ __forceinline__ void mark(int) {}
    __device__ __forceinline__ void loads_done() {}
    __device__ __forceinline__ void keep(float) {}
    __device__ __forceinline__ void task(int, int) {}
};

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE, typename Loc,
          typename Clock = NoStageClock>
__global__ __launch_bounds__(kPersistThreads) void pipelined_split_kernel(Loc loc, int* __restrict__ sync, int nsys,
                                                                          int lg_bps, int lag, Clock clk) {
    extern __shared__ float lds[];  // 10 m floats: the reduced solve
    __shared__ int s_ticket;
    constexpr int kSlices = kPersistThreads / L;
    constexpr bool kTwoLevel = METHOD == 0 && MODE == kFast && L == 16 && E >= 2;
    const int bps = 1 << lg_bps;
    const int D = pipeline_d(lag, nsys << lg_bps, bps);
    const int g = pipeline_gap(D, bps, nsys);
    const int ntk = pipeline_tickets(nsys, lg_bps);
    const int l = threadIdx.x % L;
    const int jb = threadIdx.x / L;  // slice inside the block
    int pend_q = -1;                 // system of the previous rows task, its arrival not yet counted (uniform)
    // counts the previous rows task's arrival once every wave's row stores
    // are complete (device-coherent, visible to the whole chip)
    auto arrive = [&]() {
        if (pend_q >= 0) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_fetch_add(sync + (size_t)kSplitSyncStride * (pend_q + 1), 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            pend_q = -1;
        }
    };
    if (threadIdx.x == 0) s_ticket = atomicAdd(&sync[0], 1);
    __syncthreads();
    int tk = s_ticket;
    while (tk < ntk) {  // uniform
        __syncthreads();  // every thread has read s_ticket; lds is free again
        clk.mark(0);
        int nxt = ntk;
        if (threadIdx.x == 0) nxt = atomicAdd(&sync[0], 1);
        const PipeTask task = pipeline_task(tk, nsys, lg_bps, D, g);
        const int kind = task.kind, i = task.i;
        const int q = i >> lg_bps;
        int* sys_sync = sync + (size_t)kSplitSyncStride * (q + 1);
        int* flag = sys_sync + kSplitSyncStride / 2;
        if (kind == kPipeReduced) {
            arrive();
            if (threadIdx.x == 0 && !wait_count(sys_sync, bps))
                __hip_atomic_fetch_add(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            clk.mark(1);
            reduced_rows_solve<METHOD, true>(loc.Rows(q), loc.Xed(q), loc.m, lds, threadIdx.x, kPersistThreads);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            clk.mark(2);
            if (threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            clk.mark(3);
        } else {
            const int j = (i & (bps - 1)) * kSlices + jb;  // slice inside the system
            const size_t lo = (size_t)j * (E * L) + (size_t)l * E;
            float* rows = loc.Rows(q);
            float* row = rows + (size_t)j * kRowWords;
            float* x = loc.X(q);
            int seen = 0;  // the flag, read ahead of the loads (thread 0)
            if (kind == kPipeStage3 && threadIdx.x == 0)
                seen = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            RawPart<E> raw;
            load_raw_t<E, CT>(loc.A(q), HAS_B ? loc.B(q) : nullptr, loc.C(q), loc.D(q), lo, lo, raw);
            float xo[E];
            if (kind == kPipeStage3) {
                if constexpr (REFINE) load_e<E>(x + lo, xo);
            }
            arrive();
            clk.loads_done();
            clk.mark(1);
            if constexpr (kTwoLevel && !REFINE) {
                // stage 1 needs nothing from other workgroups: one body for
                // both kinds, run before a stage-3 task waits for its flag
                Fast16State<E> fs;
                fast16_stage1<E>(raw, HAS_B, l, fs);
                clk.keep(fs.eA), clk.keep(fs.eC), clk.keep(fs.eR), clk.keep(fs.sA), clk.keep(fs.sC), clk.keep(fs.sR);
                clk.mark(2);
                if (kind == kPipeRows) {
                    if (l == 0) {
                        st_dev(row + 0, fs.sA);
                        st_dev(row + 1, fs.sC);
                        st_dev(row + 2, fs.sR);
                    }
                    if (l == L - 1) {
                        st_dev(row + 3, fs.eA);
                        st_dev(row + 4, fs.eC);
                        st_dev(row + 5, fs.eR);
                    }
                    pend_q = q;
                    clk.mark(3);
                } else {
                    if (threadIdx.x == 0 && !seen && !wait_flag(flag))
                        __hip_atomic_fetch_add(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __syncthreads();
                    clk.mark(3);
                    const float* xed = loc.Xed(q);
                    const float xe = ld_dev(xed + j);
                    const float xprev = (j > 0) ? ld_dev(xed + j - 1) : 0.0f;
                    clk.keep(xe), clk.keep(xprev);
                    clk.mark(4);
                    const int src = (int)(threadIdx.x & ~(unsigned)(L - 1)) + (L - 1);  // the slice's last lane
                    SliceEnd se;
                    se.a = __shfl(fs.eA, src, 64);
                    se.c = __shfl(fs.eC, src, 64);
                    se.r = __shfl(fs.eR, src, 64);
                    float X[E];
                    fast16_stage3<E>(fs, xprev, xe, se, l, X);
                    store_e<E>(x + lo, X);
                }
            } else if (kind == kPipeRows) {
                if constexpr (REFINE) {
                    // no x of this system is stored before its flag, which
                    // comes after every rows task of the system has loaded x
                    float xl, xr;
                    load_x_with_halo<E>(x, 0, lo, loc.n, xo, xl, xr);
                    residual_fp64<E, HAS_B>(xo, xl, xr, raw);
                    if (l == 0) st_dev(row + 6, xo[0]);
                    if (l == L - 1) st_dev(row + 7, xo[E - 1]);
                }
                float st_a, st_c, st_r, ed_a, ed_c, ed_r;
                if constexpr (kTwoLevel) {
                    Fast16State<E> fs;
                    fast16_stage1<E>(raw, HAS_B, l, fs);
                    st_a = fs.sA, st_c = fs.sC, st_r = fs.sR;
                    ed_a = fs.eA, ed_c = fs.eC, ed_r = fs.eR;
                } else {
                    float A[E], C[E], R[E];
                    normalise<E, MODE>(raw, HAS_B, A, C, R);
                    tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
                    st_a = A[0], st_c = C[0], st_r = R[0];
                    ed_a = A[E - 1], ed_c = C[E - 1], ed_r = R[E - 1];
                }
                if (l == 0) {
                    st_dev(row + 0, st_a);
                    st_dev(row + 1, st_c);
                    st_dev(row + 2, st_r);
                }
                if (l == L - 1) {
                    st_dev(row + 3, ed_a);
                    st_dev(row + 4, ed_c);
                    st_dev(row + 5, ed_r);
                }
                pend_q = q;
            } else {
                if (threadIdx.x == 0 && !seen && !wait_flag(flag))
                    __hip_atomic_fetch_add(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                if constexpr (REFINE) {
                    // inside the slice the neighbours belong to this wave,
                    // whose loads precede its stores; across slice edges, the
                    // values the rows tasks saved before any x was stored
                    const int m = loc.m;
                    float xl, xr;
                    if (l > 0) xl = x[lo - 1];
                    else xl = (j > 0) ? ld_dev(rows + (size_t)(j - 1) * kRowWords + 7) : 0.0f;
                    if (l < L - 1) xr = x[lo + E];
                    else xr = (j < m - 1) ? ld_dev(rows + (size_t)(j + 1) * kRowWords + 6) : 0.0f;
                    residual_fp64<E, HAS_B>(xo, xl, xr, raw);
                }
                const float* xed = loc.Xed(q);
                const float xe = ld_dev(xed + j);
                const float xprev = (j > 0) ? ld_dev(xed + j - 1) : 0.0f;
                const int src = (int)(threadIdx.x & ~(unsigned)(L - 1)) + (L - 1);  // the slice's last lane
                float X[E];
                if constexpr (kTwoLevel) {
                    Fast16State<E> fs;
                    fast16_stage1<E>(raw, HAS_B, l, fs);
                    SliceEnd se;
                    se.a = __shfl(fs.eA, src, 64);
                    se.c = __shfl(fs.eC, src, 64);
                    se.r = __shfl(fs.eR, src, 64);
                    fast16_stage3<E>(fs, xprev, xe, se, l, X);
                } else {
                    float A[E], C[E], R[E];
                    normalise<E, MODE>(raw, HAS_B, A, C, R);
                    if constexpr (METHOD == 0) {
                        ptpr_stage1<E, L, MODE>(A, C, R, l);
                        SliceEnd se;
                        se.a = __shfl(A[E - 1], src, 64);
                        se.c = __shfl(C[E - 1], src, 64);
                        se.r = __shfl(R[E - 1], src, 64);
                        ptpr_stage3<E, L, MODE>(A, C, R, xprev, xe, se, l, X);
                    } else {
                        float Ao[E], Co[E], Ro[E];
#pragma unroll
                        for (int c = 0; c < E; c++) {
                            Ao[c] = A[c];
                            Co[c] = C[c];
                            Ro[c] = R[c];
                        }
                        tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
                        tpr_stage3<E, L, MODE>(A, C, R, Ao, Co, Ro, xprev, xe, l, X);
                    }
                }
                if constexpr (REFINE) {
#pragma unroll
                    for (int c = 0; c < E; c++) X[c] = xo[c] + X[c];
                }
                store_e<E>(x + lo, X);
            }
        }
        if (threadIdx.x == 0) s_ticket = nxt;
        __syncthreads();
        clk.mark(5);
        clk.task(tk, kind);
        tk = s_ticket;
    }
    arrive();
}
}  // namespace tprh
