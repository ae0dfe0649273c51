// stage_clock.hip -- a diagnostic build (libtpr_diag.so, not part of the
// solver library): the batched fast-PTPR solve of one-wavefront systems on
// 16-lane slices (the shape of the 4096 x N=1024, S=256 batch), assembled
// from the same device functions as batched_kernel -- load_raw_staged
// (load_raw_t for E < 8), fast16_stage1, reduced_solve_wave, fast16_stage3,
// store_x_staged (store_e for E < 8) -- with the GPU's
// constant 100 MHz clock (s_memrealtime) read between them.  Each wave
// records six stamps: start, loads complete, stage 1 done, reduced system
// done, stage 3 done, stores complete.  The waits that make "loads complete"
// and "stores complete" exact (s_waitcnt), and the scheduling barriers and
// empty uses of each stage's results that keep the stage between its stamps,
// are the only differences from the shipped kernel; answers are the shipped kernel's bit for bit (checked by
// tpr_amd/stage_clock.py).
#include <hip/hip_runtime.h>

#include "../csrc/slice_solvers.hpp"
#include "../csrc/split_kernels.hpp"
#include "../csrc/staged_io.hpp"

namespace tprh {

constexpr int kClockBlock = 256;  // as batched_kernel for systems of one wavefront

// Makes v's values inputs of an instruction that cannot move, so that the
// stage producing them completes before the next stamp (the compiler would
// otherwise sink arithmetic whose only use is later).
__device__ __forceinline__ void keep(float v) { asm volatile("" ::"v"(v)); }
template <int N>
__device__ __forceinline__ void keep(const float (&v)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++) keep(v[i]);
}

__device__ __forceinline__ unsigned long long stamp() {
    __builtin_amdgcn_sched_barrier(0);
    const unsigned long long t = __builtin_amdgcn_s_memrealtime();
    __builtin_amdgcn_sched_barrier(0);
    return t;
}

template <int E>
__global__ __launch_bounds__(kClockBlock) void ptpr_stage_clock_kernel(const float* __restrict__ a,
                                                                       const float* __restrict__ b,
                                                                       const float* __restrict__ c,
                                                                       const float* __restrict__ d,
                                                                       float* __restrict__ x, int n, int m,
                                                                       int nsys, int lg_lanes,
                                                                       unsigned long long* __restrict__ stamps) {
    extern __shared__ float lds[];
    unsigned long long t[6];
    t[0] = stamp();
    const int sl = threadIdx.x >> lg_lanes;
    const int tt = threadIdx.x & ((1 << lg_lanes) - 1);
    float* my = lds + sl * lds_words_per_system(m);
    const int sys = blockIdx.x * (blockDim.x >> lg_lanes) + sl;
    const bool valid = sys < nsys;
    const size_t x0 = valid ? (size_t)sys * (size_t)n : 0;
    const size_t off = x0 + (size_t)tt * E;
    // the wavefront's run of 64 E floats and its LDS, as batched_kernel's
    float* stage = lds + stage_base(blockDim.x >> lg_lanes, m) + (threadIdx.x >> 6) * stage_wave_words<E>();
    const size_t wave_off = ((size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) * E;
    const size_t total = (size_t)nsys * (size_t)n;
    RawPart<E> raw;
    if constexpr (kStageX<E>) load_raw_staged<E, true>(a, b, c, d, wave_off, total, stage, threadIdx.x & 63, raw);
    else load_raw_t<E, float>(a, b, c, d, off, off, raw);
    __builtin_amdgcn_s_waitcnt(0);
    t[1] = stamp();
    const int l = tt & 15;
    const int j = tt >> 4;
    Fast16State<E> st;
    fast16_stage1<E>(raw, true, l, st);
    keep(st.y);
    keep(st.pc);
    keep(st.q);
    keep(st.eA), keep(st.eC), keep(st.eR), keep(st.sA), keep(st.sC), keep(st.sR);
    t[2] = stamp();
    float xprev, xe;
    SliceEnd se;
    reduced_solve_wave<0, 16>(m, j, (int)(threadIdx.x & 63) - tt, st.eA, st.eC, st.eR, st.sA, st.sC, st.sR, xe, xprev, se);
    keep(xprev), keep(xe), keep(se.a), keep(se.c), keep(se.r);
    t[3] = stamp();
    float X[E];
    fast16_stage3<E>(st, xprev, xe, se, l, X);
    keep(X);
    t[4] = stamp();
    if constexpr (kStageX<E>) store_x_staged<E, true>(x, X, stage, wave_off, total, threadIdx.x & 63);
    else if (valid) store_e<E, true>(x + off, X);
    __builtin_amdgcn_s_waitcnt(0);
    t[5] = stamp();
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) stamps[(size_t)wave * 6 + k] = t[k];
    }
}


// The multi-workgroup path's pipelined form: the shipped
// pipelined_split_kernel itself, instantiated for fast PTPR on 16 x 16
// slices with a diagonal, fp32, without refinement, with a clock that reads
// the GPU's 100 MHz counter at the kernel's marks (see NoStageClock in
// split_kernels.hpp) and waits for the loads before mark(1).  Per ticket,
// 8 words: the six stamps (a mark a task does not pass keeps the value of
// the last one it passed) and kind (0 rows, 1 reduced, 2 stage 3) | workgroup << 8.
struct SplitStampClock {
    unsigned long long* out;
    unsigned long long t[6];
    __device__ __forceinline__ void mark(int k) {
        const unsigned long long v = stamp();
#pragma unroll
        for (int c = 0; c < 6; c++)
            if (c >= k) t[c] = v;
    }
    __device__ __forceinline__ void loads_done() { __builtin_amdgcn_s_waitcnt(0); }
    __device__ __forceinline__ void keep(float v) { asm volatile("" ::"v"(v)); }
    __device__ __forceinline__ void task(int tk, int kind) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < 6; c++) out[(size_t)tk * 8 + c] = t[c];
            out[(size_t)tk * 8 + 6] = (unsigned long long)kind | ((unsigned long long)blockIdx.x << 8);
        }
    }
};

constexpr auto split_clock_kernel =
    pipelined_split_kernel<0, 16, 16, kFast, true, float, false, BatchLoc<float>, SplitStampClock>;
}  // namespace tprh

// The diagnostic solve: nsys systems of n = m * 16 * E equations with a
// diagonal, E = s / 16 for s = 32 .. 256 and m * 16 <= 64 lanes per system;
// stamps receives 6 values per wavefront (grid = ceil(nsys / (256 / (m*16)))
// workgroups of 4 wavefronts).  Returns a hipError_t.
extern "C" int tprh_diag_ptpr_stage_clock(const float* a, const float* b, const float* c, const float* d, float* x,
                                          int n, int s, int nsys, unsigned long long* stamps, void* stream) {
    using namespace tprh;
    const int E = s / 16, m = n / s;
    const int T = m * 16;
    if (s < 32 || s > 256 || (s & (s - 1)) || n % s || T > 64 || (m & (m - 1))) return (int)hipErrorInvalidValue;
    const int lg = __builtin_ctz((unsigned)T);
    const int spb = kClockBlock / T;
    const int grid = (nsys + spb - 1) / spb;
    size_t shmem = sizeof(float) * (size_t)spb * lds_words_per_system(m);
    if (E == 8 || E == 16)  // the staging area, as batched_kernel's
        shmem = sizeof(float) * ((size_t)stage_base(spb, m) +
                                 (size_t)(E == 8 ? stage_words<8>(kClockBlock) : stage_words<16>(kClockBlock)));
    hipStream_t st = (hipStream_t)stream;
    switch (E) {
        case 2: hipLaunchKernelGGL(ptpr_stage_clock_kernel<2>, dim3(grid), dim3(kClockBlock), shmem, st, a, b, c, d, x, n, m, nsys, lg, stamps); break;
        case 4: hipLaunchKernelGGL(ptpr_stage_clock_kernel<4>, dim3(grid), dim3(kClockBlock), shmem, st, a, b, c, d, x, n, m, nsys, lg, stamps); break;
        case 8: hipLaunchKernelGGL(ptpr_stage_clock_kernel<8>, dim3(grid), dim3(kClockBlock), shmem, st, a, b, c, d, x, n, m, nsys, lg, stamps); break;
        default: hipLaunchKernelGGL(ptpr_stage_clock_kernel<16>, dim3(grid), dim3(kClockBlock), shmem, st, a, b, c, d, x, n, m, nsys, lg, stamps); break;
    }
    return (int)hipGetLastError();
}

// The pipelined split solve of nsys fast-PTPR systems of n equations, s = 256
// (16 x 16 slices, m = n / 256 slices per system), with a diagonal;
// workspace as tprh_workspace_bytes gives for chunk = nsys; lag in rows tasks
// (0: the library's, twice the workgroups the device holds at once); the
// launch is persistent (the workgroups the device holds at once).  stamps
// receives 8 words per ticket, tprh_diag_split_tickets(n, nsys) of them.
// Returns a hipError_t.
extern "C" long long tprh_diag_split_tickets(int n, int nsys) {
    return tprh::pipeline_tickets(nsys, __builtin_ctz((unsigned)(n / 4096)));
}

extern "C" int tprh_diag_split_clock(const float* a, const float* b, const float* c, const float* d, float* x, int n,
                                     int nsys, void* workspace, int lag, unsigned long long* stamps, void* stream) {
    using namespace tprh;
    constexpr int E = 16, L = 16;
    const int m = n / (E * L);
    if (n % (E * L) || m * L < 2 * kPersistThreads || (m & (m - 1)) || m > kSplitMaxM || nsys < 1)
        return (int)hipErrorInvalidValue;
    SplitArgs p{};
    p.n = n, p.nsys = nsys, p.chunk = nsys;
    split_carve(workspace, m, nsys, &p);
    BatchLoc<float> loc;
    loc.a = a, loc.b = b, loc.c = c, loc.d = d, loc.x = x;
    loc.rows = p.rows, loc.xed = p.xed, loc.n = n, loc.m = m, loc.cstride = (size_t)n;
    const int bps = m * L / kPersistThreads;
    const int lg_bps = __builtin_ctz((unsigned)bps);
    const size_t lds = sizeof(float) * 10 * (size_t)m;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(split_clock_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(float) * 10 * kSplitMaxM));
    if (e != hipSuccess) return (int)e;
    int dev = 0, cus = 0, occ = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return (int)e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return (int)e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, split_clock_kernel, kPersistThreads, lds)) !=
        hipSuccess)
        return (int)e;
    const int cap = cus * occ;
    if (cap < 1) return (int)hipErrorInvalidValue;
    if (lag <= 0) lag = 2 * cap;
    hipStream_t st = (hipStream_t)stream;
    const int words = kSplitSyncStride * (nsys + 1);
    if ((e = hipMemsetAsync(p.sync, 0, sizeof(int) * (size_t)words, st)) != hipSuccess) return (int)e;
    SplitStampClock clk;
    clk.out = stamps;
    hipLaunchKernelGGL(split_clock_kernel, dim3(cap), dim3(kPersistThreads), lds, st, loc, p.sync, nsys, lg_bps, lag,
                       clk);
    return (int)hipGetLastError();
}
