// solve_slot.hpp -- one system solved by one group of lanes of a 1024-lane
// workgroup, from per-system base pointers.  Shared by the persistent kernels
// (uneven batches, serving batches) that put systems of different sizes and
// layouts into one launch.
#pragma once
#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "slice_solvers.hpp"

namespace tprh {

// The workgroup's lanes are cut into groups of 2^lg_lanes = m * L lanes, one
// system each (slot = threadIdx.x >> lg_lanes); the system has n = m * E * L
// equations at pa/pb/pc/pd (coefficients, right-hand side) and its answer
// goes to px.  valid == false: the slot has no system of its own, its
// pointers name another valid system, it computes but never stores.
// Every lane of the workgroup must call this (the reduced solve has
// workgroup barriers); m is uniform across the workgroup.
template <int METHOD, int MODE, bool HAS_B, int E, int L>
__device__ __forceinline__ void solve_slot(int m, int lg_lanes, bool valid, const float* __restrict__ pa,
                                           const float* __restrict__ pb, const float* __restrict__ pc,
                                           const float* __restrict__ pd, float* __restrict__ px, float* lds) {
    const int slot = threadIdx.x >> lg_lanes;
    const int t = threadIdx.x & ((1 << lg_lanes) - 1);
    const size_t off = (size_t)t * E;
    RawPart<E> raw;
    load_raw<E>(pa, HAS_B ? pb : nullptr, pc, pd, off, raw);
    float X[E];
    float* my = lds + slot * lds_words_per_system(m);
    if constexpr (METHOD == 0) {
        ptpr_solve_raw<E, L, MODE>(raw, HAS_B, m, t, my, X);
    } else {
        float A[E], C[E], R[E];
        normalise<E, MODE>(raw, HAS_B, A, C, R);
        tpr_solve_regs<E, L, MODE>(A, C, R, m, t, my, X);
    }
    if (valid) store_e<E>(px + off, X);
}

}  // namespace tprh
