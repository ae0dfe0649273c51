// split_impl.hpp -- PTPR / TPR for systems larger than one workgroup: the
// launch code of the split path, instantiated per method and arithmetic by
// split_ptpr_fast.hip, split_ptpr_exact.hip, split_tpr_fast.hip and
// split_tpr_exact.hip (compiled in parallel); split.hip dispatches.
//
// The reference needs every block of a system co-resident for its grid-wide
// barriers and limits n to s^2 (src/ptpr.cu:543,581,655).  Here a batch is
// solved in rounds of `chunk` systems, each round in one of three forms:
//
// Two passes (split_kernels.hpp), the library's form:
//   K1 rows_kernel   : every lane group reduces its slices and writes two
//                      boundary rows per slice (E_st before its update, E_ed);
//   K2 reduced_kernel: one workgroup per system applies E_st += s1 E_ed
//                      (src/ptpr.cpp:193-211), builds the m reduced rows
//                      (:218-241) and solves them in LDS (PCR for PTPR,
//                      src/pcr.cpp; CR for TPR, src/cr.cpp);
//   K3 stage3_kernel : every lane group reloads its slices, redoes stage 1
//                      and runs stage 3 (src/ptpr.cpp:263-287 /
//                      src/tpr.cpp:277-308): the inputs are read twice,
//                      36 B/eq of traffic.
// K1 needs only the first and last row of every slice.  In PTPR's in-slice
// PCR the final value of a slice's first (last) row depends only on the rows
// at offsets 0 mod 2^p (-1 mod 2^p) of the previous levels, updated with the
// same formulas in the same order as TPR's tree stage 1 (src/tpr.cpp:142-204
// vs src/ptpr.cpp:140-182), so in exact arithmetic K1 runs the O(S) tree for
// both methods and its rows are bit-identical to the full PCR's.  In fast
// arithmetic, PTPR on 16-lane slices runs the two-level stage 1 in K1 and K3
// (as the one-workgroup path does), which also keeps both kernels at 72-78
// VGPRs, 6-7 waves per SIMD: on MI355X the 256 x 65536 batch takes 41 + 6 +
// 50 us this way against 42 + 6 + 58 us with the register PCR in K3.
//
// One pass (persistent_split_kernel), on request (split_form 2): each block
// keeps its slices in registers from the load to the store and waits for
// its system's reduced solve (20 B/eq).  It needs every block of a system
// resident at once, and each wave of resident blocks exposes a reduced
// solve's latency with nothing to overlap it: 0.25 ms for 256 x 65536.
//
// Pipelined (pipelined_split_kernel), on request (split_form 3): one
// persistent launch of rows, reduced and stage-3 tasks in ticket order, the
// reduced solves overlapping other systems' loads and the reloads meant to
// hit the last-level cache.  0.136-0.149 ms for 256 x 65536: its waits on
// device-coherent atomics (microseconds each behind the memory traffic) and
// its 108 VGPRs (4 waves per SIMD, so fewer bytes in flight) cost more than
// the reloads it saves, which the last-level cache serves only about 1.4x
// faster than HBM on this part.
//
// The library chooses two passes.  Exact answers are the same bits in every
// form (the same operations in the same order); fast answers agree within
// rounding (the forms differ in where stage 1's end rows come from).
#pragma once
#include <hip/hip_runtime.h>

#include "split_kernels.hpp"

namespace tprh {

namespace {

// Clears the one-pass form's counters and flags (a kernel, not a memset, so
// that a captured graph replays it correctly).
__global__ void clear_words_kernel(int* __restrict__ w, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) w[i] = 0;
}

template <typename CT>
BatchLoc<CT> round_loc(const SplitArgs& p, size_t base, int m) {
    const size_t cbase = p.shared_coef ? 0 : base;
    BatchLoc<CT> loc;
    loc.a = static_cast<const CT*>(p.a) + cbase;
    loc.b = p.b ? static_cast<const CT*>(p.b) + cbase : nullptr;
    loc.c = static_cast<const CT*>(p.c) + cbase;
    loc.d = p.d + base;
    loc.x = p.x + base;
    loc.rows = p.rows;
    loc.xed = p.xed;
    loc.n = p.n;
    loc.m = m;
    loc.cstride = p.shared_coef ? 0 : (size_t)p.n;
    return loc;
}

// Two-pass round: K1 rows, K2 one workgroup per system for the reduced
// system, K3 stage 3 (which reads the round's inputs a second time).
template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
hipError_t launch_round_two_pass(const SplitArgs& p, size_t base, int cs, int m, size_t lds, hipStream_t stream) {
    const BatchLoc<CT> loc = round_loc<CT>(p, base, m);
    const int lg_m = __builtin_ctz((unsigned)m);
    const long long nslices = (long long)cs * m;
    const int grid = (int)((nslices * L + 255) / 256);
    hipLaunchKernelGGL((rows_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>), dim3(grid), dim3(256), 0, stream,
                       loc, lg_m, nslices);
    // K2 touches only the rows and x_ed: one instantiation for every storage type
    BatchLoc<float> rl{};
    rl.rows = p.rows;
    rl.xed = p.xed;
    rl.n = p.n;
    rl.m = m;
    int rthreads = m < 1024 ? m : 1024;
    rthreads = (rthreads + 63) / 64 * 64;
    hipLaunchKernelGGL((reduced_kernel<METHOD, BatchLoc<float>>), dim3(cs), dim3(rthreads), lds, stream, rl);
    hipLaunchKernelGGL((stage3_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>), dim3(grid), dim3(256), 0,
                       stream, loc, lg_m, nslices);
    return hipGetLastError();
}

// Workgroups of the one-pass kernel a device holds at once, with `lds`
// bytes of LDS each (0 on a query error).
template <typename K>
int persistent_capacity(K kernel, size_t lds) {
    int dev = 0, cus = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kPersistThreads, lds) != hipSuccess) return 0;
    return cus * occ;
}

// Lag of the pipelined form in rows tasks: twice the workgroups resident at
// once (capacity; its query failing gives 0), so that a task's waits are
// normally over when it starts (measured on MI355X at 0.5, 1, 2 and 4 times
// the capacity: 0.149, 0.140, 0.139, 0.136 ms for 256 x 65536).
int pipeline_lag(int capacity) { return 2 * capacity; }

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
hipError_t launch_round_persistent(const SplitArgs& p, size_t base, int cs, int m, int lg_bps, int capacity,
                                   size_t lds, hipStream_t stream) {
    const BatchLoc<CT> loc = round_loc<CT>(p, base, m);
    const int words = kSplitSyncStride * (cs + 1);
    hipLaunchKernelGGL(clear_words_kernel, dim3((words + 255) / 256), dim3(256), 0, stream, p.sync, words);
    const long long blocks = (long long)cs << lg_bps;
    const int grid = (int)(blocks < capacity ? blocks : capacity);
    hipLaunchKernelGGL((persistent_split_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>), dim3(grid),
                       dim3(kPersistThreads), lds, stream, loc, p.sync, cs, lg_bps);
    return hipGetLastError();
}

// Pipelined round: one persistent launch (at most `capacity` workgroups)
// running a rows task and a stage-3 task per block of slices, taken in
// ticket order with a lag of `lag` blocks between the two.
template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
hipError_t launch_round_pipelined(const SplitArgs& p, size_t base, int cs, int m, int lg_bps, int capacity, int lag,
                                  size_t lds, hipStream_t stream) {
    const BatchLoc<CT> loc = round_loc<CT>(p, base, m);
    const int words = kSplitSyncStride * (cs + 1);
    hipLaunchKernelGGL(clear_words_kernel, dim3((words + 255) / 256), dim3(256), 0, stream, p.sync, words);
    const long long tasks = pipeline_tickets(cs, lg_bps);
    const int grid = (int)(tasks < capacity ? tasks : (capacity > 0 ? capacity : 1));
    hipLaunchKernelGGL((pipelined_split_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>), dim3(grid),
                       dim3(kPersistThreads), lds, stream, loc, p.sync, cs, lg_bps, lag, NoStageClock{});
    return hipGetLastError();
}

template <int METHOD, int E, int L, int MODE, bool HAS_B, typename CT, bool REFINE>
hipError_t launch_split_variant(const SplitArgs& p, int m, hipStream_t stream) {
    // a system is bps blocks of 256 lanes (m L > 1024 on this path, so bps
    // is a power of two >= 4)
    const int bps = (int)(((long long)m * L) / kPersistThreads);
    const int lg_bps = __builtin_ctz((unsigned)bps);
    const size_t lds = sizeof(float) * 10 * (size_t)m;
    const int lds_max = (int)(sizeof(float) * 10 * kSplitMaxM);
    const auto pk = persistent_split_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>;
    static std::atomic<unsigned long long> plds_set{0}, rlds_set{0};
    hipError_t e = ensure_lds_limit(reinterpret_cast<const void*>(pk), lds_max, plds_set);
    if (e == hipSuccess)
        e = ensure_lds_limit(reinterpret_cast<const void*>(reduced_kernel<METHOD, BatchLoc<float>>), lds_max, rlds_set);
    if (e != hipSuccess) return e;
    // the one-pass form only when the caller asks for it (measured slower
    // than two passes on MI355X wherever a round has more blocks than the
    // chip holds at once: every wave of blocks waits out a reduced solve),
    // and only where every block of a system can be resident at once
    const int capacity = p.form == 2 ? persistent_capacity(pk, lds) : 0;
    const bool one_pass = p.form == 2 && capacity >= bps;
    int lag = 0, qcap = 0;
    if (p.form == 3) {
        const auto qk = pipelined_split_kernel<METHOD, E, L, MODE, HAS_B, CT, REFINE, BatchLoc<CT>>;
        static std::atomic<unsigned long long> qlds_set{0};
        if ((e = ensure_lds_limit(reinterpret_cast<const void*>(qk), lds_max, qlds_set)) != hipSuccess) return e;
        qcap = persistent_capacity(qk, lds);
        // the lag: about the workgroups the device holds at once, so that a
        // stage-3 task is normally taken after its system's reduced solve
        lag = pipeline_lag(qcap);
    }
    for (int s0 = 0; s0 < p.nsys; s0 += p.chunk) {
        const int cs = (p.nsys - s0 < p.chunk) ? p.nsys - s0 : p.chunk;
        const size_t base = (size_t)s0 * p.n;
        if (p.form == 3)
            e = launch_round_pipelined<METHOD, E, L, MODE, HAS_B, CT, REFINE>(p, base, cs, m, lg_bps, qcap, lag, lds,
                                                                              stream);
        else if (one_pass)
            e = launch_round_persistent<METHOD, E, L, MODE, HAS_B, CT, REFINE>(p, base, cs, m, lg_bps, capacity, lds,
                                                                               stream);
        else
            e = launch_round_two_pass<METHOD, E, L, MODE, HAS_B, CT, REFINE>(p, base, cs, m, lds, stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int METHOD, int E, int L, int MODE>
hipError_t launch_split_el(const SplitArgs& p, hipStream_t stream) {
    const int m = p.n / (E * L);
    const bool hb = p.b != nullptr;
    if (p.f16) {
        return hb ? launch_split_variant<METHOD, E, L, MODE, true, _Float16, false>(p, m, stream)
                  : launch_split_variant<METHOD, E, L, MODE, false, _Float16, false>(p, m, stream);
    }
    if (p.refine) {
        return hb ? launch_split_variant<METHOD, E, L, MODE, true, float, true>(p, m, stream)
                  : launch_split_variant<METHOD, E, L, MODE, false, float, true>(p, m, stream);
    }
    return hb ? launch_split_variant<METHOD, E, L, MODE, true, float, false>(p, m, stream)
              : launch_split_variant<METHOD, E, L, MODE, false, float, false>(p, m, stream);
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

}  // namespace tprh
