// tpr_device.hpp -- device-side building blocks shared by the TPR/PTPR kernels.
//
// Lane layout used by every batched kernel in this library:
//   * a slice of S = E * L equations is owned by L consecutive lanes of one
//     wave64, each lane holding E consecutive equations in registers;
//   * a system of N = m * S equations is owned by m * L consecutive lanes.
// With L == 16 a slice is exactly one DPP row, so every cross-lane exchange of
// the in-slice reduction is a single DPP row shift whose out-of-row lanes
// receive the slice's virtual boundary value for free.
//
// Arithmetic policies:
//   Exact -- the reference's CPU operation order, no FMA, IEEE division
//            (the translation unit is built with -ffp-contract=off), so results
//            are bit-identical to src/ptpr.cpp / src/tpr.cpp on x86-64.
//   Fast  -- FMA and the hardware reciprocal (v_rcp_f32) in the hot loops.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tprh {

constexpr int kWave = 64;

enum Arith : int { kFast = 0, kExact = 1 };

__device__ __forceinline__ float as_f(int v) { return __builtin_bit_cast(float, v); }
__device__ __forceinline__ int as_i(float v) { return __builtin_bit_cast(int, v); }

// DPP controls (GFX9 encoding).
constexpr int dpp_row_shr(int d) { return 0x110 + d; }
constexpr int dpp_row_shl(int d) { return 0x100 + d; }
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

// v from lane (lane - D) of the same slice, or bnd for the first D lanes of
// the slice.  l is the lane index inside the slice (0 .. L-1).
template <int L, int D>
__device__ __forceinline__ float shift_up(float v, float bnd, int l) {
    static_assert(D >= 1, "shift distance");
    if constexpr (L == 1) {
        return bnd;
    } else if constexpr (L == 16) {
        return as_f(__builtin_amdgcn_update_dpp(as_i(bnd), as_i(v), dpp_row_shr(D), 0xF, 0xF, false));
    } else if constexpr (L < 16) {
        float t = as_f(__builtin_amdgcn_update_dpp(as_i(bnd), as_i(v), dpp_row_shr(D), 0xF, 0xF, false));
        return l >= D ? t : bnd;
    } else {
        float t;
        if constexpr (D == 1) {
            t = as_f(__builtin_amdgcn_update_dpp(as_i(bnd), as_i(v), kDppWaveShr1, 0xF, 0xF, false));
        } else {
            int src = (int)__lane_id() - D;
            t = as_f(__builtin_amdgcn_ds_bpermute(src << 2, as_i(v)));
        }
        return l >= D ? t : bnd;
    }
}

// v from lane (lane + D) of the same slice, or bnd for the last D lanes.
template <int L, int D>
__device__ __forceinline__ float shift_down(float v, float bnd, int l) {
    static_assert(D >= 1, "shift distance");
    if constexpr (L == 1) {
        return bnd;
    } else if constexpr (L == 16) {
        return as_f(__builtin_amdgcn_update_dpp(as_i(bnd), as_i(v), dpp_row_shl(D), 0xF, 0xF, false));
    } else if constexpr (L < 16) {
        float t = as_f(__builtin_amdgcn_update_dpp(as_i(bnd), as_i(v), dpp_row_shl(D), 0xF, 0xF, false));
        return l < L - D ? t : bnd;
    } else {
        float t;
        if constexpr (D == 1) {
            t = as_f(__builtin_amdgcn_update_dpp(as_i(bnd), as_i(v), kDppWaveShl1, 0xF, 0xF, false));
        } else {
            int src = (int)__lane_id() + D;
            t = as_f(__builtin_amdgcn_ds_bpermute((src & 63) << 2, as_i(v)));
        }
        return l < L - D ? t : bnd;
    }
}

// Reciprocal used by the Fast policy: v_rcp_f32 (1 ulp).
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 1 / x rounded to nearest, the same bits as IEEE division (1.0f / x) gives,
// in four instructions instead of eleven: the hardware reciprocal (within an
// ulp), one Newton step in FMA, and v_div_fixup for zeros, infinities and
// NaNs.  Checked against 1.0f / x for every one of the 2^32 float inputs by
// tests/test_gpu_exact_recip.py (tpr_amd/diag/exact_div.hip); inputs whose
// reciprocal leaves the normal range (|x| < 2^-126 or |x| > 2^126) take the
// IEEE division, behind a branch the wavefront skips when no lane needs it.
__device__ __forceinline__ float recip_rn(float x) {
    const unsigned ax = __builtin_bit_cast(unsigned, x) & 0x7fffffffu;
    if (ax - 0x00800000u >= 0x7e800000u - 0x00800000u) return 1.0f / x;
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_amdgcn_div_fixupf(__builtin_fmaf(e, r, r), x, 1.0f);
}

// The same without the range branch, for the exhaustive check of where the
// short sequence alone is exact.
__device__ __forceinline__ float recip_rn_unguarded(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_amdgcn_div_fixupf(__builtin_fmaf(e, r, r), x, 1.0f);
}

// ---------------------------------------------------------------------------
// Fused DPP forms for slices of L == 16 lanes (one DPP row), Fast policy.
//
// The fast PTPR stage 1 keeps the off-diagonals negated (A = -a, C = -c) so
// that every virtual boundary partner is either a zero (C and r of the left
// virtual row, A and r of the right one: DPP zero-fill, bound_ctrl) or a
// multiplicative identity (A = 1 of the left row, C = 1 of the right row: the
// DPP lane is disabled and the destination keeps its value).  Each exchange
// then folds into the ALU instruction that consumes it.  The compiler's DPP
// combiner does not form these (it leaves an FMA or an identity multiply
// unfused), so they are written out here.  "s_nop 1" gives the two wait
// states GFX9 requires between a VALU write of a VGPR and a DPP read of it.
// The accumulators of the multi-instruction blocks are early-clobber: they are
// written before the last input is read, so they must not share a register
// with any input (rn starts equal to r, which is also an input).

// den -= Cs(lane-D) * ak;  den -= As(lane+D) * ck;
// rn  += Rs(lane-D) * ak;  rn  += Qs(lane+D) * ck      (zero past the row ends)
template <int D>
__device__ __forceinline__ void dpp_den_rn(float& den, float& rn, float Cs, float As, float Rs, float Qs, float ak,
                                           float ck) {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %[den], -%[cs], %[ak] row_shr:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %[den], -%[as], %[ck] row_shl:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %[rn], %[rs], %[ak] row_shr:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %[rn], %[qs], %[ck] row_shl:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : [den] "+&v"(den), [rn] "+&v"(rn)
        : [cs] "v"(Cs), [as] "v"(As), [rs] "v"(Rs), [qs] "v"(Qs), [ak] "v"(ak), [ck] "v"(ck), [d] "i"(D));
}

// Left partner only: den -= Cs(lane-D) * ak;  rn += Rs(lane-D) * ak.
template <int D>
__device__ __forceinline__ void dpp_den_rn_up(float& den, float& rn, float Cs, float Rs, float ak) {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %[den], -%[cs], %[ak] row_shr:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %[rn], %[rs], %[ak] row_shr:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : [den] "+&v"(den), [rn] "+&v"(rn)
        : [cs] "v"(Cs), [rs] "v"(Rs), [ak] "v"(ak), [d] "i"(D));
}

// Right partner only: den -= As(lane+D) * ck;  rn += Rs(lane+D) * ck.
template <int D>
__device__ __forceinline__ void dpp_den_rn_dn(float& den, float& rn, float As, float Rs, float ck) {
    asm("s_nop 1\n\t"
        "v_fmac_f32_dpp %[den], -%[as], %[ck] row_shl:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
        "v_fmac_f32_dpp %[rn], %[rs], %[ck] row_shl:%[d] row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : [den] "+&v"(den), [rn] "+&v"(rn)
        : [as] "v"(As), [rs] "v"(Rs), [ck] "v"(ck), [d] "i"(D));
}

// u = As(lane-D) * u, or u unchanged in the first D lanes of the row.
template <int D>
__device__ __forceinline__ float dpp_keep_mul_up(float As, float u) {
    asm("s_nop 1\n\tv_mul_f32_dpp %[u], %[as], %[u] row_shr:%[d] row_mask:0xf bank_mask:0xf"
        : [u] "+v"(u)
        : [as] "v"(As), [d] "i"(D));
    return u;
}

// v = Cs(lane+D) * v, or v unchanged in the last D lanes of the row.
template <int D>
__device__ __forceinline__ float dpp_keep_mul_dn(float Cs, float v) {
    asm("s_nop 1\n\tv_mul_f32_dpp %[v], %[cs], %[v] row_shl:%[d] row_mask:0xf bank_mask:0xf"
        : [v] "+v"(v)
        : [cs] "v"(Cs), [d] "i"(D));
    return v;
}

// One equation of a tridiagonal system with unit diagonal:
//   a * x[k - u] + x[k] + c * x[k + u] = r.
struct Eq {
    float a, c, r;
};

// PCR-style update of E_k by its left (kl) and right (kr) partners, in the
// operation order of PTPR stage 1 (src/ptpr.cpp:140-182) and TPR stage 1
// (src/tpr.cpp:142-204).  hasL / hasR select the reference's upper / interior
// / lower formulas; in Fast mode the caller passes the virtual boundary
// partners (a=-1,c=0,r=0 on the left; a=0,c=-1,r=0 on the right), which make
// the single interior formula reproduce the other two, and the flags are
// ignored.
template <int MODE>
__device__ __forceinline__ Eq pcr_update(Eq k, Eq kl, Eq kr, bool hasL, bool hasR) {
    Eq o;
    if constexpr (MODE == kExact) {
        // inv = 1 / (1 - c[kl]*a[k] - a[kr]*c[k])
        float t1 = hasL ? (1.0f - kl.c * k.a) : 1.0f;
        float den = hasR ? (t1 - kr.a * k.c) : t1;
        float inv = recip_rn(den);
        o.a = hasL ? (-inv * kl.a) * k.a : inv * k.a;
        o.c = hasR ? (-inv * kr.c) * k.c : inv * k.c;
        float rn = hasL ? (k.r - kl.r * k.a) : k.r;
        rn = hasR ? (rn - kr.r * k.c) : rn;
        o.r = inv * rn;
    } else {
        // the operation order of the fused DPP form (slice_solvers.hpp,
        // ptpr_stage1_fast16): left partner first, inv times the row's own
        // coefficient before the partner's, so that every slice shape gives
        // the same bits
        float den = __builtin_fmaf(-kr.a, k.c, __builtin_fmaf(-kl.c, k.a, 1.0f));
        float inv = fast_rcp(den);
        o.a = (-inv * k.a) * kl.a;
        o.c = (-inv * k.c) * kr.c;
        float rn = __builtin_fmaf(-kr.r, k.c, __builtin_fmaf(-kl.r, k.a, k.r));
        o.r = inv * rn;
    }
    return o;
}

// ---------------------------------------------------------------------------
// Synchronisation of the lanes that share one system's LDS rows.  WAVE: the
// system lives inside one wavefront, whose LDS instructions execute in
// program order, so a compiler fence at wavefront scope orders them and no
// workgroup barrier is needed; otherwise a workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void group_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Reduced-system solvers on LDS (exact reference operation order).
// Rows live in LDS arrays a[], c[], r[] of length m (a power of two); the
// calling threads t = 0 .. nthr-1 cooperate; every thread of the group
// (the wavefront when WAVE, else the workgroup) must call these.

// PCR, src/pcr.cpp:15-89.  Double-buffered in (a, c, r) and (a1, c1, r1);
// returns a pointer to the buffer that holds the answer in its r part.
template <bool WAVE = false>
__device__ __forceinline__ float* lds_pcr(float* a, float* c, float* r, float* a1, float* c1,
                                          float* r1, int m, int t, int nthr) {
    int pn = 31 - __builtin_clz(m);
    for (int p = 0; p < pn; p++) {
        int s = 1 << p;
        for (int k = t; k < m; k += nthr) {
            float e, na, nc, nr;
            if (k < s) {
                int kr = k + s;
                e = recip_rn(1.0f - c[k] * a[kr]);
                na = e * a[k];
                nc = -e * c[k] * c[kr];
                nr = e * (r[k] - c[k] * r[kr]);
            } else if (k < m - s) {
                int kl = k - s, kr = k + s;
                float ap = a[k], cp = c[k];
                e = recip_rn(1.0f - ap * c[kl] - cp * a[kr]);
                na = -e * ap * a[kl];
                nc = -e * cp * c[kr];
                nr = e * (r[k] - ap * r[kl] - cp * r[kr]);
            } else {
                int kl = k - s;
                float ap = a[k], cp = c[k];
                e = recip_rn(1.0f - ap * c[kl]);
                na = -e * ap * a[kl];
                nc = e * cp;
                nr = e * (r[k] - ap * r[kl]);
            }
            a1[k] = na;
            c1[k] = nc;
            r1[k] = nr;
        }
        group_sync<WAVE>();
        float* t0 = a; a = a1; a1 = t0;
        t0 = c; c = c1; c1 = t0;
        t0 = r; r = r1; r1 = t0;
    }
    return r;
}

}  // namespace tprh

namespace tprh {

// CR, src/cr.cpp:23-116, over m rows (a, c, r) in LDS; the answer goes to
// xo[0..m).  Threads t = 0 .. nthr-1 cooperate; every thread of the
// workgroup must call it.
template <bool WAVE = false>
__device__ __forceinline__ void lds_cr(float* a, float* c, float* r, float* xo, int m, int t, int nthr) {
    if (m <= 1) {
        if (t == 0) xo[0] = r[0];
        group_sync<WAVE>();
        return;
    }
    const int lg = 31 - __builtin_clz(m);
    // forward reduction, src/cr.cpp:38-82: rows i = ux-1, 2ux-1, ... ; the
    // last row uses the lower formula.  Values of a level are computed into
    // registers first and written after a barrier (the reference's aa/cc/rr).
    for (int p = 0; p < lg - 1; p++) {
        const int u = 1 << p;
        const int ux = 1 << (p + 1);
        const int rows = m / ux;
        for (int q0 = 0; q0 < rows; q0 += nthr) {
            const int q = q0 + t;
            const int i = ux - 1 + q * ux;
            float na = 0.0f, nc = 0.0f, nr = 0.0f;
            const bool act = q < rows;
            if (act) {
                if (i < m - ux) {
                    int kl = i - u, k = i, kr = i + u;
                    float inv = recip_rn(1.0f - c[kl] * a[k] - a[kr] * c[k]);
                    na = -inv * a[kl] * a[k];
                    nc = -inv * c[kr] * c[k];
                    nr = inv * (r[k] - r[kl] * a[k] - r[kr] * c[k]);
                } else {
                    int k = m - 1, kl = k - u;
                    float inv = recip_rn(1.0f - c[kl] * a[k]);
                    na = -inv * a[kl] * a[k];
                    nc = inv * c[k];
                    nr = inv * (r[k] - r[kl] * a[k]);
                }
            }
            group_sync<WAVE>();
            if (act) {
                a[i] = na;
                c[i] = nc;
                r[i] = nr;
            }
            group_sync<WAVE>();
        }
    }
    // backward substitution, src/cr.cpp:88-116
    if (t == 0) {
        int i = m / 2 - 1, u = m / 2;
        float inv_det = recip_rn(1.0f - c[i] * a[i + u]);
        xo[i] = (r[i] - c[i] * r[i + u]) * inv_det;
        xo[i + u] = (r[i + u] - r[i] * a[i + u]) * inv_det;
    }
    group_sync<WAVE>();
    for (int k = lg - 2; k >= 0; k--) {
        const int u = 1 << k;
        const int rows = m / (2 * u);
        for (int q = t; q < rows; q += nthr) {
            const int i = u - 1 + q * 2 * u;
            if (q == 0) {
                xo[i] = r[i] - c[i] * xo[i + u];
            } else if (i < m - u) {
                xo[i] = r[i] - a[i] * xo[i - u] - c[i] * xo[i + u];
            }
        }
        group_sync<WAVE>();
    }
}

}  // namespace tprh
