// staged_io.hpp -- a wavefront's loads and stores of 64 lanes x E
// consecutive floats as 1 KB runs per instruction.
//
// Every kernel here gives lane t of a system the equations [t E, t E + E),
// so the 64 lanes of a wavefront own one contiguous run of 64 E floats of
// each array.  Accessed lane by lane, each 16-byte load or store instruction
// touches a quarter (E = 16) of 64 separate 64-byte pieces; the non-temporal
// stores then reach memory as partial lines (WRITE_SIZE 1.5 times the bytes
// of x on 4096 x N=1024).  Passed through 68 E floats of LDS per wavefront,
// every instruction covers 1 KB of consecutive bytes instead, whole lines:
// 4096 x N=1024 batched PTPR 18.6 -> 14.3-14.8 us.
#pragma once
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "slice_solvers.hpp"

namespace tprh {

// The shapes that stage (16-byte pieces of the lanes' runs).
template <int E>
constexpr bool kStageX = E == 8 || E == 16;

// The 16-byte pieces of a wavefront's run sit in LDS with one spare slot
// after every 16 (slot q + q / 16): the per-lane reads and writes of 4 or 2
// consecutive pieces, like the per-instruction ones of 64 consecutive
// pieces, then fall on 16 different 16-byte bank groups in every group of
// 16 lanes, free of bank conflicts.
__device__ __forceinline__ int stage_slot(int q) { return q + (q >> 4); }

// LDS words per wavefront, and per workgroup of `lanes` lanes, for the
// staging: 17 E / 16 slots of 16 bytes per lane.
template <int E>
__host__ __device__ constexpr int stage_wave_words() { return 68 * E; }
template <int E>
__host__ __device__ constexpr int stage_words(int lanes) { return (lanes / 64) * stage_wave_words<E>(); }

// Start of the staging area behind spb systems' rows of m slices (16-byte
// aligned).
__host__ __device__ constexpr int stage_base(int spb, int m) { return (spb * lds_words_per_system(m) + 3) & ~3; }

typedef float stage_v4f __attribute__((ext_vector_type(4)));

// stage: the wavefront's stage_wave_words<E>() floats of LDS.
// Stores X of the 64 lanes to x[run .. run + 64 E): the wavefront's run.
// Pieces at or past `total` are not stored (lanes of systems past the end of
// a batch).  Every lane of the wavefront calls this.  NT: non-temporal.
template <int E, bool NT>
__device__ __forceinline__ void store_x_staged(float* __restrict__ x, const float (&X)[E], float* stage, size_t run,
                                               size_t total, int lane) {
    stage_v4f* s4 = reinterpret_cast<stage_v4f*>(stage);
#pragma unroll
    for (int i = 0; i < E / 4; i++)
        s4[stage_slot(lane * (E / 4) + i)] = stage_v4f{X[4 * i], X[4 * i + 1], X[4 * i + 2], X[4 * i + 3]};
    group_sync<true>();
#pragma unroll
    for (int k = 0; k < E / 4; k++) {
        const int q = k * 64 + lane;  // 16-byte piece of the run
        const size_t gi = run + 4 * (size_t)q;
        const stage_v4f v = s4[stage_slot(q)];
        if (gi < total) {
            if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<stage_v4f*>(x + gi));
            else *reinterpret_cast<stage_v4f*>(x + gi) = v;
        }
    }
    group_sync<true>();  // the area may be written again
}

// Loads the lane's E values of a, b (HAS_B), c and d from the wavefront's
// run at `run`, as 1 KB per instruction.  Pieces at or past `total` load the
// last piece before it instead (lanes that never store).  Every lane of the
// wavefront calls this.
template <int E, bool HAS_B>
__device__ __forceinline__ void load_raw_staged(const float* __restrict__ a, const float* __restrict__ b,
                                                const float* __restrict__ c, const float* __restrict__ d, size_t run,
                                                size_t total, float* stage, int lane, RawPart<E>& p) {
    constexpr int K = E / 4;
    size_t gi[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const size_t g = run + 4 * (size_t)(k * 64 + lane);
        gi[k] = g < total ? g : total - 4;
    }
    stage_v4f la[K], lb[K], lc[K], ld[K];
#pragma unroll
    for (int k = 0; k < K; k++) la[k] = *reinterpret_cast<const stage_v4f*>(a + gi[k]);
#pragma unroll
    for (int k = 0; k < K; k++) lc[k] = *reinterpret_cast<const stage_v4f*>(c + gi[k]);
#pragma unroll
    for (int k = 0; k < K; k++) ld[k] = *reinterpret_cast<const stage_v4f*>(d + gi[k]);
    if constexpr (HAS_B) {
#pragma unroll
        for (int k = 0; k < K; k++) lb[k] = *reinterpret_cast<const stage_v4f*>(b + gi[k]);
    }
    stage_v4f* s4 = reinterpret_cast<stage_v4f*>(stage);
    auto turn = [&](const stage_v4f(&l)[K], float(&out)[E]) {
#pragma unroll
        for (int k = 0; k < K; k++) s4[stage_slot(k * 64 + lane)] = l[k];
        group_sync<true>();
#pragma unroll
        for (int i = 0; i < K; i++) {
            const stage_v4f v = s4[stage_slot(lane * K + i)];
            out[4 * i] = v.x;
            out[4 * i + 1] = v.y;
            out[4 * i + 2] = v.z;
            out[4 * i + 3] = v.w;
        }
        group_sync<true>();
    };
    turn(la, p.a);
    turn(lc, p.c);
    turn(ld, p.d);
    if constexpr (HAS_B) turn(lb, p.b);
}

}  // namespace tprh
