// slice_solvers.hpp -- the PTPR and TPR solve of one system by a group of
// lanes, as device functions shared by the batched, uneven-batch and
// multi-RHS kernels.
//
// Data flow of one solve (all in one kernel, no grid-wide synchronisation):
//   1. each lane loads E consecutive equations of a/b/c/d (vector loads) and
//      divides by the diagonal in registers (the 4-argument setter,
//      src/ptpr.hpp:80-90);
//   2. stage 1 in registers: PTPR runs the in-slice PCR of src/ptpr.cpp:130-190,
//      TPR the tree reduction of src/tpr.cpp:123-230; partners in other lanes
//      arrive by DPP row shifts (L == 16) or ds_bpermute (L > 16);
//   3. the E_st += s1 * E_ed update (src/ptpr.cpp:193-211) and the reduced
//      system (src/ptpr.cpp:218-241) go through LDS, one row per slice;
//   4. stage 2 solves the m-row reduced system in LDS: PCR for PTPR
//      (src/pcr.cpp:15-89), CR for TPR (src/cr.cpp:23-116);
//   5. stage 3 in registers: PTPR's one-step substitution (src/ptpr.cpp:263-287)
//      or TPR's tree substitution (src/tpr.cpp:277-308);
//   6. each lane stores its E answers.
// HBM traffic is the algorithmic minimum: 16 B/eq read (12 with a unit
// diagonal) and 4 B/eq written.
#pragma once
#include <float.h>

#include "launch.hpp"
#include "tpr_device.hpp"

namespace tprh {

// ---------------------------------------------------------------------------
// Loads and stores of E consecutive floats.
template <int E>
__device__ __forceinline__ void load_e(const float* __restrict__ p, float (&v)[E]) {
    if constexpr (E % 4 == 0) {
#pragma unroll
        for (int i = 0; i < E / 4; i++) {
            float4 t = reinterpret_cast<const float4*>(p)[i];
            v[4 * i + 0] = t.x;
            v[4 * i + 1] = t.y;
            v[4 * i + 2] = t.z;
            v[4 * i + 3] = t.w;
        }
    } else if constexpr (E == 2) {
        float2 t = reinterpret_cast<const float2*>(p)[0];
        v[0] = t.x;
        v[1] = t.y;
    } else {
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = p[i];
    }
}

// NT: non-temporal stores (marked streaming for the caches).  On MI355X
// they make the one-workgroup batched kernel's x stores cheaper (4096 x
// N=1024: PTPR 18.8-19.0 -> 18.65-18.72 us, TPR 19.6-19.9 -> 18.8-18.9 us)
// and the persistent uneven and multi-workgroup kernels' much dearer
// (0.382 -> 0.452 ms, 0.098 -> 0.122 ms), so only the batched kernel asks.
template <int E, bool NT = false>
__device__ __forceinline__ void store_e(float* __restrict__ p, const float (&v)[E]) {
    if constexpr (E % 4 == 0 && NT) {
        typedef float v4f __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int i = 0; i < E / 4; i++) {
            const v4f t = {v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]};
            __builtin_nontemporal_store(t, reinterpret_cast<v4f*>(p) + i);
        }
    } else if constexpr (E % 4 == 0) {
#pragma unroll
        for (int i = 0; i < E / 4; i++) {
            reinterpret_cast<float4*>(p)[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
        }
    } else if constexpr (E == 2) {
        reinterpret_cast<float2*>(p)[0] = make_float2(v[0], v[1]);
    } else {
#pragma unroll
        for (int i = 0; i < E; i++) p[i] = v[i];
    }
}

// One lane's E equations as stored in memory (before the division by the
// diagonal).
template <int E>
struct RawPart {
    float a[E], b[E], c[E], d[E];
};

// Branch-free: a lane of a system past the end of the batch loads the same
// lane of system 0 (always present) and never stores, so the loads stay in
// one basic block and the compiler keeps them where they are issued.
template <int E>
__device__ __forceinline__ void load_raw(const float* __restrict__ a, const float* __restrict__ b,
                                         const float* __restrict__ c, const float* __restrict__ d, size_t off,
                                         RawPart<E>& p) {
    load_e<E>(a + off, p.a);
    load_e<E>(c + off, p.c);
    load_e<E>(d + off, p.d);
    if (b != nullptr) load_e<E>(b + off, p.b);
}

// E consecutive fp16 values, widened to fp32 (exact).
template <int E>
__device__ __forceinline__ void load_e(const _Float16* __restrict__ p, float (&v)[E]) {
    if constexpr (E % 8 == 0) {
#pragma unroll
        for (int i = 0; i < E / 8; i++) {
            uint4 t = reinterpret_cast<const uint4*>(p)[i];
            const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                v[8 * i + 2 * k] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w[k] & 0xFFFF));
                v[8 * i + 2 * k + 1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w[k] >> 16));
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = (float)p[i];
    }
}

// Coefficients stored as CT (float or _Float16), right-hand side as float.
// coff: offset of the lane's coefficients, doff: of its right-hand side
// (they differ when several right-hand sides share one matrix).
template <int E, typename CT>
__device__ __forceinline__ void load_raw_t(const CT* __restrict__ a, const CT* __restrict__ b,
                                           const CT* __restrict__ c, const float* __restrict__ d, size_t coff,
                                           size_t doff, RawPart<E>& p) {
    load_e<E>(a + coff, p.a);
    load_e<E>(c + coff, p.c);
    load_e<E>(d + doff, p.d);
    if (b != nullptr) load_e<E>(b + coff, p.b);
}

// Iterative refinement: replaces p.d by the residual r = d - A x of the
// system at the current x, accumulated in FP64 in the order
// ((d - b x) - a x_left) - c x_right (products of two floats are exact in
// double, so the only roundings are the three subtractions and the final
// conversion).  xo holds the lane's own E values of x; xl / xr are x just
// before / after them (0 outside the system: a[0] and c[n-1] are outside the
// matrix).  HAS_B == false: unit diagonal.
template <int E, bool HAS_B>
__device__ __forceinline__ void residual_fp64(const float (&xo)[E], float xl, float xr, RawPart<E>& p) {
#pragma unroll
    for (int k = 0; k < E; k++) {
        const double xm = (k == 0) ? (double)xl : (double)xo[k - 1];
        const double xp = (k == E - 1) ? (double)xr : (double)xo[k + 1];
        const double bk = HAS_B ? (double)p.b[k] : 1.0;
        double acc = (double)p.d[k] - bk * (double)xo[k];
        acc -= (double)p.a[k] * xm;
        acc -= (double)p.c[k] * xp;
        p.d[k] = (float)acc;
    }
}

// Loads the lane's x values and their outer neighbours from memory (x0: the
// system's first equation, i0: the lane's first equation).
template <int E>
__device__ __forceinline__ void load_x_with_halo(const float* __restrict__ x, size_t x0, size_t i0, int n,
                                                 float (&xo)[E], float& xl, float& xr) {
    load_e<E>(x + i0, xo);
    xl = (i0 > x0) ? x[i0 - 1] : 0.0f;
    xr = (i0 + E < x0 + (size_t)n) ? x[i0 + E] : 0.0f;
}

// Divides by the diagonal as the 4-argument setter does (src/ptpr.hpp:80-90);
// has_b == false means a unit diagonal (the reference's GPU entry points).
template <int E, int MODE>
__device__ __forceinline__ void normalise(const RawPart<E>& p, bool has_b, float (&A)[E], float (&C)[E],
                                          float (&R)[E]) {
#pragma unroll
    for (int i = 0; i < E; i++) {
        if (has_b) {
            if constexpr (MODE == kExact) {
                A[i] = p.a[i] / p.b[i];
                C[i] = p.c[i] / p.b[i];
                R[i] = p.d[i] / p.b[i];
            } else {
                const float rb = fast_rcp(p.b[i]);
                A[i] = p.a[i] * rb;
                C[i] = p.c[i] * rb;
                R[i] = p.d[i] * rb;
            }
        } else {
            A[i] = p.a[i];
            C[i] = p.c[i];
            R[i] = p.d[i];
        }
    }
}

// ---------------------------------------------------------------------------
// PTPR stage 1: in-slice PCR, levels u = U, 2U, ... < S (src/ptpr.cpp:134-190).
template <int E, int L, int MODE, int U>
__device__ __forceinline__ void ptpr_stage1_level(float (&A)[E], float (&C)[E], float (&R)[E], int l) {
    constexpr int S = E * L;
    if constexpr (U < E) {
        // Partners inside the lane, except the first / last U equations whose
        // partners sit in the previous / next lane of the slice.
        Eq left[U], right[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            left[j].a = shift_up<L, 1>(A[E - U + j], -1.0f, l);
            left[j].c = shift_up<L, 1>(C[E - U + j], 0.0f, l);
            left[j].r = shift_up<L, 1>(R[E - U + j], 0.0f, l);
            right[j].a = shift_down<L, 1>(A[j], 0.0f, l);
            right[j].c = shift_down<L, 1>(C[j], -1.0f, l);
            right[j].r = shift_down<L, 1>(R[j], 0.0f, l);
        }
        float nA[E], nC[E], nR[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            Eq k{A[j], C[j], R[j]};
            Eq kl = (j >= U) ? Eq{A[j - U], C[j - U], R[j - U]} : left[j];
            Eq kr = (j + U < E) ? Eq{A[j + U], C[j + U], R[j + U]} : right[j + U - E];
            bool hasL = (j >= U) || (l > 0);
            bool hasR = (j + U < E) || (l < L - 1);
            Eq o = pcr_update<MODE>(k, kl, kr, hasL, hasR);
            nA[j] = o.a;
            nC[j] = o.c;
            nR[j] = o.r;
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            A[j] = nA[j];
            C[j] = nC[j];
            R[j] = nR[j];
        }
    } else {
        // Partner of register j is register j of the lane D lanes away.
        constexpr int D = U / E;
        bool hasL = l >= D;
        bool hasR = l < L - D;
#pragma unroll
        for (int j = 0; j < E; j++) {
            Eq kl{shift_up<L, D>(A[j], -1.0f, l), shift_up<L, D>(C[j], 0.0f, l), shift_up<L, D>(R[j], 0.0f, l)};
            Eq kr{shift_down<L, D>(A[j], 0.0f, l), shift_down<L, D>(C[j], -1.0f, l),
                  shift_down<L, D>(R[j], 0.0f, l)};
            Eq o = pcr_update<MODE>(Eq{A[j], C[j], R[j]}, kl, kr, hasL, hasR);
            A[j] = o.a;
            C[j] = o.c;
            R[j] = o.r;
        }
    }
    if constexpr (2 * U < S) ptpr_stage1_level<E, L, MODE, 2 * U>(A, C, R, l);
}

// ---------------------------------------------------------------------------
// PTPR stage 1, Fast policy, slices of one DPP row (L == 16).  Ap = -a and
// Cp = -c (negated off-diagonals); every cross-lane partner is folded into a
// fused DPP instruction (see dpp_den_rn / dpp_keep_mul_*).  In the negated
// form the PCR update of src/ptpr.cpp:155-168 reads
//   den = 1 - C'kl A'k - A'kr C'k,  inv = 1/den,
//   A'  = A'kl (inv A'k),  C' = C'kr (inv C'k),  r = inv (rk + rkl A'k + rkr C'k),
// and the upper / lower formulas (:140-152, :171-182) are the same expression
// with the missing partner's C'kl, rkl (resp. A'kr, rkr) zero and A'kl
// (resp. C'kr) one.
template <int E, int U>
__device__ __forceinline__ void ptpr_stage1_fast16(float (&Ap)[E], float (&Cp)[E], float (&R)[E]) {
    constexpr int S = E * 16;
    if constexpr (U < E) {
        float nA[E], nC[E], nR[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            const bool lx = j < U;       // left partner in the previous lane (register E-U+j)
            const bool rx = j + U >= E;  // right partner in the next lane (register j+U-E)
            float ak = Ap[j], ck = Cp[j];
            float den, rn;
            if (!lx && !rx) {
                den = __builtin_fmaf(-Ap[j + U], ck, __builtin_fmaf(-Cp[j - U], ak, 1.0f));
                rn = __builtin_fmaf(R[j + U], ck, __builtin_fmaf(R[j - U], ak, R[j]));
            } else if (lx) {
                // left partner first, as in the other branches
                den = 1.0f;
                rn = R[j];
                dpp_den_rn_up<1>(den, rn, Cp[lx ? E - U + j : 0], R[lx ? E - U + j : 0], ak);
                den = __builtin_fmaf(-Ap[j + U], ck, den);
                rn = __builtin_fmaf(R[j + U], ck, rn);
            } else {
                den = __builtin_fmaf(-Cp[j - U], ak, 1.0f);
                rn = __builtin_fmaf(R[j - U], ak, R[j]);
                dpp_den_rn_dn<1>(den, rn, Ap[rx ? j + U - E : 0], R[rx ? j + U - E : 0], ck);
            }
            const float inv = fast_rcp(den);
            float u = inv * ak, v = inv * ck;
            nA[j] = lx ? dpp_keep_mul_up<1>(Ap[lx ? E - U + j : 0], u) : Ap[lx ? 0 : j - U] * u;
            nC[j] = rx ? dpp_keep_mul_dn<1>(Cp[rx ? j + U - E : 0], v) : Cp[rx ? 0 : j + U] * v;
            nR[j] = inv * rn;
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            Ap[j] = nA[j];
            Cp[j] = nC[j];
            R[j] = nR[j];
        }
    } else {
        constexpr int D = U / E;
#pragma unroll
        for (int j = 0; j < E; j++) {
            const float ak = Ap[j], ck = Cp[j];
            float den = 1.0f, rn = R[j];
            dpp_den_rn<D>(den, rn, ck, ak, R[j], R[j], ak, ck);
            const float inv = fast_rcp(den);
            const float nA = dpp_keep_mul_up<D>(ak, inv * ak);
            const float nC = dpp_keep_mul_dn<D>(ck, inv * ck);
            R[j] = inv * rn;
            Ap[j] = nA;
            Cp[j] = nC;
        }
    }
    if constexpr (2 * U < S) ptpr_stage1_fast16<E, 2 * U>(Ap, Cp, R);
}

// ---------------------------------------------------------------------------
// TPR stage 1: the tree reduction (src/tpr.cpp:142-216).  At level p (block
// size P = 2^p, partner distance U = P/2) only block starts (offset % P == 0)
// and block ends (offset % P == P-1) change.  Even-offset equations keep their
// original values in Ao/Co/Ro for stage 3 (the reference's bkup of
// src/tpr.cpp:128-132); odd offsets end with their stage-1 values
// (src/tpr.cpp:209-213).
template <int E, int L, int MODE, int U>
__device__ __forceinline__ void tpr_stage1_level(float (&A)[E], float (&C)[E], float (&R)[E], int l) {
    constexpr int S = E * L;
    constexpr int P = 2 * U;
    if constexpr (P <= E) {
        // Blocks lie inside one lane.  The first block start needs register
        // E-U of the previous lane; the last block end needs register U-1 of
        // the next lane.
        Eq left{shift_up<L, 1>(A[E - U], -1.0f, l), shift_up<L, 1>(C[E - U], 0.0f, l),
                shift_up<L, 1>(R[E - U], 0.0f, l)};
        Eq right{shift_down<L, 1>(A[U - 1], 0.0f, l), shift_down<L, 1>(C[U - 1], -1.0f, l),
                 shift_down<L, 1>(R[U - 1], 0.0f, l)};
        float nA[E], nC[E], nR[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            nA[j] = A[j];
            nC[j] = C[j];
            nR[j] = R[j];
        }
#pragma unroll
        for (int j = 0; j < E; j += P) {
            // block start j
            {
                Eq kl = (j >= U) ? Eq{A[j - U], C[j - U], R[j - U]} : left;
                Eq kr{A[j + U], C[j + U], R[j + U]};
                bool hasL = (j >= U) || (l > 0);
                Eq o = pcr_update<MODE>(Eq{A[j], C[j], R[j]}, kl, kr, hasL, true);
                nA[j] = o.a;
                nC[j] = o.c;
                nR[j] = o.r;
            }
            // block end j + P - 1
            {
                const int k = j + P - 1;
                Eq kl{A[k - U], C[k - U], R[k - U]};
                Eq kr = (k + U < E) ? Eq{A[k + U], C[k + U], R[k + U]} : right;
                bool hasR = (k + U < E) || (l < L - 1);
                Eq o = pcr_update<MODE>(Eq{A[k], C[k], R[k]}, kl, kr, true, hasR);
                nA[k] = o.a;
                nC[k] = o.c;
                nR[k] = o.r;
            }
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            A[j] = nA[j];
            C[j] = nC[j];
            R[j] = nR[j];
        }
    } else {
        // Blocks span Q = P/E lanes.  Starts are register 0 of lanes with
        // l % Q == 0, ends register E-1 of lanes with l % Q == Q-1; partners
        // sit D = U/E lanes away in the same register.
        constexpr int D = U / E;
        constexpr int Q = P / E;
        const int lq = l & (Q - 1);
        bool hasL = l >= D;
        bool hasR = l < L - D;
        Eq s0{A[0], C[0], R[0]};
        Eq sl{shift_up<L, D>(s0.a, -1.0f, l), shift_up<L, D>(s0.c, 0.0f, l), shift_up<L, D>(s0.r, 0.0f, l)};
        Eq sr{shift_down<L, D>(s0.a, 0.0f, l), shift_down<L, D>(s0.c, -1.0f, l),
              shift_down<L, D>(s0.r, 0.0f, l)};
        Eq e0{A[E - 1], C[E - 1], R[E - 1]};
        Eq el{shift_up<L, D>(e0.a, -1.0f, l), shift_up<L, D>(e0.c, 0.0f, l), shift_up<L, D>(e0.r, 0.0f, l)};
        Eq er{shift_down<L, D>(e0.a, 0.0f, l), shift_down<L, D>(e0.c, -1.0f, l),
              shift_down<L, D>(e0.r, 0.0f, l)};
        Eq so = pcr_update<MODE>(s0, sl, sr, hasL, hasR);
        Eq eo = pcr_update<MODE>(e0, el, er, hasL, hasR);
        if (lq == 0) {
            A[0] = so.a;
            C[0] = so.c;
            R[0] = so.r;
        }
        if (lq == Q - 1) {
            A[E - 1] = eo.a;
            C[E - 1] = eo.c;
            R[E - 1] = eo.r;
        }
    }
    if constexpr (P < S) tpr_stage1_level<E, L, MODE, P>(A, C, R, l);
}

// TPR stage 3 (src/tpr.cpp:277-308): tree back-substitution.  X[E-1] of the
// slice's last lane holds x_ed on entry; xprev is x[st-1] (0 for slice 0).
// Coefficients: even offsets use the originals (Ao...), odd offsets the
// stage-1 values (A...).
template <int E, int L, int MODE, int U>
__device__ __forceinline__ void tpr_stage3_level(const float (&A)[E], const float (&C)[E], const float (&R)[E],
                                                 const float (&Ao)[E], const float (&Co)[E],
                                                 const float (&Ro)[E], float (&X)[E], float xprev, int l) {
    auto coef = [&](int j, float& ca, float& cc, float& cr) {
        // offset parity: E even -> parity of j; E == 1 -> parity of l
        bool even = (E % 2 == 0) ? ((j & 1) == 0) : ((l & 1) == 0);
        ca = even ? Ao[j] : A[j];
        cc = even ? Co[j] : C[j];
        cr = even ? Ro[j] : R[j];
    };
    auto upd = [&](float ca, float cc, float cr, float xl, float xr) -> float {
        if constexpr (MODE == kExact) {
            return cr - ca * xl - cc * xr;
        } else {
            return __builtin_fmaf(-cc, xr, __builtin_fmaf(-ca, xl, cr));
        }
    };
    if constexpr (U >= E) {
        constexpr int D = U / E;
        float xl = shift_up<L, D>(X[E - 1], xprev, l);
        float xr = shift_down<L, D>(X[E - 1], 0.0f, l);
        if (((l + 1) & (2 * D - 1)) == D) {
            float ca, cc, cr;
            coef(E - 1, ca, cc, cr);
            X[E - 1] = upd(ca, cc, cr, xl, xr);
        }
        if constexpr (U > 1) tpr_stage3_level<E, L, MODE, U / 2>(A, C, R, Ao, Co, Ro, X, xprev, l);
    } else {
        // In-lane levels U = E/2 .. 1; x of the previous lane's last equation
        // is final by now.
        float xleft = shift_up<L, 1>(X[E - 1], xprev, l);
#pragma unroll
        for (int u = E / 2; u >= 1; u /= 2) {
#pragma unroll
            for (int j = u - 1; j < E; j += 2 * u) {
                float xl = (j >= u) ? X[j - u] : xleft;
                float ca, cc, cr;
                coef(j, ca, cc, cr);
                X[j] = upd(ca, cc, cr, xl, X[j + u]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The middle of the solve, shared by PTPR and TPR: E_st update, reduced
// system, stage 2.  On return xr[j] holds x at the end of slice j.
// lds: lds_words_per_system(m) floats for this system.
// (eA, eC, eR) is E_ed in slice lane L-1, (sA, sC, sR) is E_st in slice lane 0.
template <int METHOD, bool WAVE = false>  // 0 = PTPR (PCR stage 2), 1 = TPR (CR stage 2)
__device__ __forceinline__ const float* reduced_solve(float* lds, int m, int j, int l, int L, int t,
                                                      float eA, float eC, float eR, float sA, float sC,
                                                      float sR) {
    float* ed_a = lds;
    float* ed_c = ed_a + m;
    float* ed_r = ed_c + m;
    float* st_a = ed_r + m;
    float* st_c = st_a + m;
    float* st_r = st_c + m;
    float* ra0 = st_r + m;
    float* rc0 = ra0 + m;
    float* rr0 = rc0 + m;
    float* ra1 = rr0 + m;
    float* rc1 = ra1 + m;
    float* rr1 = rc1 + m;
    float* xo = rr1 + m;

    if (l == L - 1) {
        ed_a[j] = eA;
        ed_c[j] = eC;
        ed_r[j] = eR;
    }
    group_sync<WAVE>();
    if (l == 0) {
        // E_st += s1 * E_ed, src/ptpr.cpp:193-211 (src/tpr.cpp:327-340)
        float c_ed = ed_c[j];
        float s1 = (fabsf(c_ed) < FLT_EPSILON) ? -1.0f : -sC / c_ed;
        st_a[j] = sA + s1 * ed_a[j];
        st_c[j] = s1;
        st_r[j] = sR + s1 * ed_r[j];
    }
    group_sync<WAVE>();
    if (t < m) {
        // reduced row t from (E_ed(t), E_st(t+1)), src/ptpr.cpp:218-241
        if (t < m - 1) {
            float ak = ed_a[t], ck = ed_c[t], rk = ed_r[t];
            float akr = st_a[t + 1], ckr = st_c[t + 1], rkr = st_r[t + 1];
            float inv = recip_rn(1.0f - akr * ck);
            ra0[t] = inv * ak;
            rc0[t] = -inv * ckr * ck;
            rr0[t] = inv * (rk - rkr * ck);
        } else {
            ra0[t] = ed_a[t];
            rc0[t] = ed_c[t];
            rr0[t] = ed_r[t];
        }
    }
    group_sync<WAVE>();
    if constexpr (METHOD == 0) {
        return lds_pcr<WAVE>(ra0, rc0, rr0, ra1, rc1, rr1, m, t, m);
    } else {
        lds_cr<WAVE>(ra0, rc0, rr0, xo, m, t, m);
        return xo;
    }
}

// Scratch of one system's middle step, read back by stage 3.
struct SliceEnd {
    float a, c, r;  // E_ed after stage 1
};

__device__ __forceinline__ SliceEnd slice_end(const float* lds, int m, int j) {
    return SliceEnd{lds[j], lds[m + j], lds[2 * m + j]};
}

// reduced_solve for a system whose m slices of L lanes lie in one
// wavefront (m L <= 64), in its lanes base .. base + m L - 1: the same
// operations in the same order (so the same bits), exchanged by cross-lane
// reads (ds_bpermute) instead of LDS stores, barriers and loads -- for
// m = 4 a chain of five exchanges instead of five store-barrier-load rounds,
// which shortens the tail of a batch whose last wavefronts finish their
// loads last.  Row j of the reduced system lives in the lanes of slice j.
// Every lane gets x at the end of its slice (xe) and of the slice before
// (xprev, 0 for the first) and its slice's E_ed (se).  Every lane of the
// wavefront calls it.
template <int METHOD, int L>  // 0 = PTPR (PCR stage 2), 1 = TPR (CR stage 2)
__device__ __forceinline__ void reduced_solve_wave(int m, int j, int base, float eA, float eC, float eR, float sA,
                                                   float sC, float sR, float& xe, float& xprev, SliceEnd& se) {
    auto lane = [&](int k) { return base + L * k; };  // a lane of slice k
    const int me = lane(j);
    const float ed_a = __shfl(eA, me + L - 1, 64), ed_c = __shfl(eC, me + L - 1, 64),
                ed_r = __shfl(eR, me + L - 1, 64);
    se = SliceEnd{ed_a, ed_c, ed_r};
    const float s_a = __shfl(sA, me, 64), s_c = __shfl(sC, me, 64), s_r = __shfl(sR, me, 64);
    // E_st += s1 * E_ed, src/ptpr.cpp:193-211 (src/tpr.cpp:327-340)
    const float s1 = (fabsf(ed_c) < FLT_EPSILON) ? -1.0f : -s_c / ed_c;
    const float st_a = s_a + s1 * ed_a, st_c = s1, st_r = s_r + s1 * ed_r;
    // reduced row j from (E_ed(j), E_st(j+1)), src/ptpr.cpp:218-241
    const int nx = lane(j + 1 < m ? j + 1 : j);
    const float akr = __shfl(st_a, nx, 64), ckr = __shfl(st_c, nx, 64), rkr = __shfl(st_r, nx, 64);
    float a, c, r;
    if (j < m - 1) {
        const float inv = recip_rn(1.0f - akr * ed_c);
        a = inv * ed_a;
        c = -inv * ckr * ed_c;
        r = inv * (ed_r - rkr * ed_c);
    } else {
        a = ed_a;
        c = ed_c;
        r = ed_r;
    }
    if constexpr (METHOD == 0) {
        // PCR over the m rows, lds_pcr's operations (src/pcr.cpp:15-89)
        for (int s = 1; s < m; s <<= 1) {
            const int kl = lane(j >= s ? j - s : j);
            const int kr = lane(j + s < m ? j + s : j);
            const float al = __shfl(a, kl, 64), cl = __shfl(c, kl, 64), rl = __shfl(r, kl, 64);
            const float ar = __shfl(a, kr, 64), cr = __shfl(c, kr, 64), rr = __shfl(r, kr, 64);
            float e, na, nc, nr;
            if (j < s) {
                e = recip_rn(1.0f - c * ar);
                na = e * a;
                nc = -e * c * cr;
                nr = e * (r - c * rr);
            } else if (j < m - s) {
                const float ap = a, cp = c;
                e = recip_rn(1.0f - ap * cl - cp * ar);
                na = -e * ap * al;
                nc = -e * cp * cr;
                nr = e * (r - ap * rl - cp * rr);
            } else {
                const float ap = a, cp = c;
                e = recip_rn(1.0f - ap * cl);
                na = -e * ap * al;
                nc = e * cp;
                nr = e * (r - ap * rl);
            }
            a = na;
            c = nc;
            r = nr;
        }
        xe = r;
    } else {
        // CR over the m rows, lds_cr's operations (src/cr.cpp:23-116)
        if (m <= 1) {
            xe = r;
        } else {
            // forward reduction: at level u, rows i = 2u-1, 4u-1, ... take
            // rows i - u and i + u; the last row the lower formula
            for (int u = 1; 2 * u < m; u <<= 1) {
                const int kl = lane(j >= u ? j - u : j);
                const int kr = lane(j + u < m ? j + u : j);
                const float al = __shfl(a, kl, 64), cl = __shfl(c, kl, 64), rl = __shfl(r, kl, 64);
                const float ar = __shfl(a, kr, 64), cr = __shfl(c, kr, 64), rr = __shfl(r, kr, 64);
                if (((j + 1) & (2 * u - 1)) == 0) {
                    float na, nc, nr;
                    if (j < m - 2 * u) {
                        const float inv = recip_rn(1.0f - cl * a - ar * c);
                        na = -inv * al * a;
                        nc = -inv * cr * c;
                        nr = inv * (r - rl * a - rr * c);
                    } else {
                        const float inv = recip_rn(1.0f - cl * a);
                        na = -inv * al * a;
                        nc = inv * c;
                        nr = inv * (r - rl * a);
                    }
                    a = na;
                    c = nc;
                    r = nr;
                }
            }
            // backward substitution: the 2 x 2 system of rows m/2-1 and m-1,
            // then each level's rows from the ones above it
            const int h = m / 2;
            const int top = lane(h - 1), bot = lane(m - 1);
            const float c_t = __shfl(c, top, 64), r_t = __shfl(r, top, 64);
            const float a_b = __shfl(a, bot, 64), r_b = __shfl(r, bot, 64);
            const float inv_det = recip_rn(1.0f - c_t * a_b);
            float xo = 0.0f;
            if (j == h - 1) xo = (r_t - c_t * r_b) * inv_det;
            if (j == m - 1) xo = (r_b - r_t * a_b) * inv_det;
            for (int u = h / 2; u >= 1; u >>= 1) {
                const float xl = __shfl(xo, lane(j >= u ? j - u : j), 64);
                const float xr = __shfl(xo, lane(j + u < m ? j + u : j), 64);
                if (((j + 1) & (2 * u - 1)) == u) {  // rows u-1, 3u-1, ...
                    if (j == u - 1) xo = r - c * xr;
                    else if (j < m - u) xo = r - a * xl - c * xr;
                }
            }
            xe = xo;
        }
    }
    xprev = __shfl(xe, lane(j > 0 ? j - 1 : 0), 64);
    if (j == 0) xprev = 0.0f;
}

// ---------------------------------------------------------------------------
// Stage entry points shared by the batched (one workgroup per system) and the
// split (several workgroups per system) kernels.

// PTPR stage 1 on the lane's registers (src/ptpr.cpp:130-190).
template <int E, int L, int MODE>
__device__ __forceinline__ void ptpr_stage1(float (&A)[E], float (&C)[E], float (&R)[E], int l) {
    if constexpr (MODE == kFast && L == 16) {
#pragma unroll
        for (int i = 0; i < E; i++) {
            A[i] = -A[i];
            C[i] = -C[i];
        }
        ptpr_stage1_fast16<E, 1>(A, C, R);
#pragma unroll
        for (int i = 0; i < E; i++) {
            A[i] = -A[i];
            C[i] = -C[i];
        }
    } else {
        ptpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
    }
}

// PTPR stage 3 (src/ptpr.cpp:263-287): x_i = r_i - a_i x[st-1] - c_i key with
// key = (r_ed - a_ed x[st-1] - x_ed) / c_ed (0 if c_ed == 0); the slice's last
// equation takes x_ed.  A/C/R hold the stage-1 values (E_st before its update).
template <int E, int L, int MODE>
__device__ __forceinline__ void ptpr_stage3(const float (&A)[E], const float (&C)[E], const float (&R)[E],
                                            float xprev, float xe, SliceEnd se, int l, float (&X)[E]) {
    float key;
    if constexpr (MODE == kExact) {
        key = recip_rn(se.c) * (se.r - se.a * xprev - xe);
    } else {
        key = fast_rcp(se.c) * (__builtin_fmaf(-se.a, xprev, se.r) - xe);
    }
    if (se.c == 0.0f) key = 0.0f;
#pragma unroll
    for (int i = 0; i < E; i++) {
        if constexpr (MODE == kExact) {
            X[i] = R[i] - A[i] * xprev - C[i] * key;
        } else {
            X[i] = __builtin_fmaf(-C[i], key, __builtin_fmaf(-A[i], xprev, R[i]));
        }
    }
    if (l == L - 1) X[E - 1] = xe;
}

// TPR stage 3 (src/tpr.cpp:277-308) from the stage-3 coefficients (A/C/R:
// stage-1 values, Ao/Co/Ro: originals; even offsets use the originals).
template <int E, int L, int MODE>
__device__ __forceinline__ void tpr_stage3(const float (&A)[E], const float (&C)[E], const float (&R)[E],
                                           const float (&Ao)[E], const float (&Co)[E], const float (&Ro)[E],
                                           float xprev, float xe, int l, float (&X)[E]) {
#pragma unroll
    for (int i = 0; i < E; i++) X[i] = 0.0f;
    X[E - 1] = (l == L - 1) ? xe : 0.0f;
    tpr_stage3_level<E, L, MODE, (E * L) / 2>(A, C, R, Ao, Co, Ro, X, xprev, l);
}

// ---------------------------------------------------------------------------
// Whole PTPR solve of one system by its lanes, from the raw loads of the
// lane's E equations at element offset `off`.  sys_valid == false lanes run
// the same instruction stream on zeros (every lane must take part in the
// cross-lane exchanges and the barriers) and store nothing.
// WAVE: the system's lanes lie inside one wavefront (m * L <= 64).
template <int E, int L, int MODE, bool WAVE = false>
__device__ __forceinline__ void ptpr_solve_regs(float (&A)[E], float (&C)[E], float (&R)[E], int m, int t,
                                                float* lds, float (&X)[E]) {
    const int l = t % L;
    const int j = t / L;
    ptpr_stage1<E, L, MODE>(A, C, R, l);
    if constexpr (WAVE) {
        // the system's lanes are t = 0 .. m L - 1 of one wavefront
        float xe, xprev;
        SliceEnd se;
        reduced_solve_wave<0, L>(m, j, (int)(threadIdx.x & 63) - t, A[E - 1], C[E - 1], R[E - 1], A[0], C[0], R[0], xe,
                              xprev, se);
        ptpr_stage3<E, L, MODE>(A, C, R, xprev, xe, se, l, X);
    } else {
        const float* xr = reduced_solve<0, WAVE>(lds, m, j, l, L, t, A[E - 1], C[E - 1], R[E - 1], A[0], C[0], R[0]);
        const float xprev = (j > 0) ? xr[j - 1] : 0.0f;
        ptpr_stage3<E, L, MODE>(A, C, R, xprev, xr[j], slice_end(lds, m, j), l, X);
    }
}

template <int E, int L, int MODE, bool WAVE = false>
__device__ __forceinline__ void tpr_solve_regs(float (&A)[E], float (&C)[E], float (&R)[E], int m, int t,
                                               float* lds, float (&X)[E]) {
    const int l = t % L;
    const int j = t / L;
    float Ao[E], Co[E], Ro[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        Ao[i] = A[i];
        Co[i] = C[i];
        Ro[i] = R[i];
    }
    tpr_stage1_level<E, L, MODE, 1>(A, C, R, l);
    if constexpr (WAVE) {
        // the system's lanes are t = 0 .. m L - 1 of one wavefront
        float xe, xprev;
        SliceEnd se;
        reduced_solve_wave<1, L>(m, j, (int)(threadIdx.x & 63) - t, A[E - 1], C[E - 1], R[E - 1], A[0], C[0], R[0], xe,
                                 xprev, se);
        tpr_stage3<E, L, MODE>(A, C, R, Ao, Co, Ro, xprev, xe, l, X);
    } else {
        const float* xr = reduced_solve<1, WAVE>(lds, m, j, l, L, t, A[E - 1], C[E - 1], R[E - 1], A[0], C[0], R[0]);
        const float xprev = (j > 0) ? xr[j - 1] : 0.0f;
        tpr_stage3<E, L, MODE>(A, C, R, Ao, Co, Ro, xprev, xr[j], l, X);
    }
}

// ---------------------------------------------------------------------------
// PTPR, Fast policy, slices of one DPP row (L == 16) with E >= 2 equations per
// lane.  Stage 1's result -- every row of the slice as
//   x_i = R_i - A_i x[st-1] - C_i x[ed+1]
// -- is the same linear map whatever order eliminates the slice's interior;
// here it is reached in two levels instead of log2(S) PCR levels over every
// equation:
//   (1) in each lane, the first E-1 rows are eliminated by a Thomas sweep
//       (forward elimination, back-substitution of three right-hand sides):
//         x_k = y_k + p_k e_{l-1} + q_k e_l,   k = 0 .. E-2,
//       where e_l is the lane's last unknown and e_{l-1} the previous
//       lane's (x[st-1] for the slice's first lane); the division by the
//       diagonal is part of the sweep;
//   (2) each lane's last row, with its neighbours' x_k substituted, is one
//       row of a tridiagonal system in the e's (x[st-1] and x[ed+1] at its
//       ends), the slice's 16 lanes one row each; PTPR's own stage 1 on
//       that system -- the in-slice PCR levels at lane distances 1, 2, 4, 8,
//       through the fused DPP forms -- leaves e_l = R_l - A_l x[st-1] -
//       C_l x[ed+1].
// The slice's last row is lane 15's e, its first row follows from lane 0's
// x_0; the reduced system (src/ptpr.cpp:193-258) is solved as for every
// other shape.  Stage 3 (src/ptpr.cpp:263-287) computes key = x[ed+1] from
// the slice's last row as the reference does, then every lane its e_l and
// its rows x_k = y_k + p_k e_{l-1} + q_k e_l.
// Work per lane: about 9 VALU and one reciprocal per equation in (1), and
// four cross-lane levels of one row in (2), against log2(S) levels of 9 VALU
// and a reciprocal per equation in the PCR form.  Mathematically the same
// stage 1; rounding differs from the reference's order, which the exact
// policy keeps.
// The lane's state between the two halves of the two-level solve.
template <int E>
struct Fast16State {
    float y[E - 1], pc[E - 1], q[E - 1];  // x_k = y_k + pc_k e_{l-1} + q_k e_l
    float eA, eC, eR;                     // e_l = eR - eA x[st-1] - eC x[ed+1]
    float sA, sC, sR;                     // the slice's first row (valid in lane 0)
};

// First half: stages (1) and (2) above, from the raw loads.
template <int E>
__device__ __forceinline__ void fast16_stage1(const RawPart<E>& p, bool has_b, int l, Fast16State<E>& st) {
    static_assert(E >= 2, "the lane needs at least one interior row");
    // (1) forward elimination over rows 0 .. E-2: c' in cp, y' in y, p' in pc
    float cp[E - 1];
    float(&y)[E - 1] = st.y;
    float(&pc)[E - 1] = st.pc;
    float(&q)[E - 1] = st.q;
    {
        const float inv = fast_rcp(has_b ? p.b[0] : 1.0f);
        cp[0] = p.c[0] * inv;
        y[0] = p.d[0] * inv;
        pc[0] = -p.a[0] * inv;
    }
#pragma unroll
    for (int k = 1; k < E - 1; k++) {
        const float den = __builtin_fmaf(-p.a[k], cp[k - 1], has_b ? p.b[k] : 1.0f);
        const float inv = fast_rcp(den);
        cp[k] = p.c[k] * inv;
        y[k] = __builtin_fmaf(-p.a[k], y[k - 1], p.d[k]) * inv;
        pc[k] = (-p.a[k] * pc[k - 1]) * inv;
    }
    // back-substitution: y_k, p_k (coefficient of e_{l-1}), q_k (of e_l)
    q[E - 2] = -cp[E - 2];
#pragma unroll
    for (int k = E - 3; k >= 0; k--) {
        y[k] = __builtin_fmaf(-cp[k], y[k + 1], y[k]);
        pc[k] = __builtin_fmaf(-cp[k], pc[k + 1], pc[k]);
        q[k] = -cp[k] * q[k + 1];
    }
    // (2) the lane's last row: a x_{E-2} + b e_l + c x_E = d with x_E the next
    // lane's x_0 = y0' + p0' e_l + q0' e_{l+1}; past the slice's last lane
    // x_E is x[ed+1] itself (y0' = p0' = 0, q0' = 1)
    const float yn = shift_down<16, 1>(y[0], 0.0f, l);
    const float pn = shift_down<16, 1>(pc[0], 0.0f, l);
    const float qn = shift_down<16, 1>(q[0], 1.0f, l);
    const float aL = p.a[E - 1], cL = p.c[E - 1];
    const float beta = __builtin_fmaf(cL, pn, __builtin_fmaf(aL, q[E - 2], has_b ? p.b[E - 1] : 1.0f));
    const float delta = __builtin_fmaf(-cL, yn, __builtin_fmaf(-aL, y[E - 2], p.d[E - 1]));
    const float ib = fast_rcp(beta);
    // negated off-diagonals, the form the fused DPP levels take
    float Ap[1] = {-((aL * pc[E - 2]) * ib)};
    float Cp[1] = {-((cL * qn) * ib)};
    float Rr[1] = {delta * ib};
    ptpr_stage1_fast16<1, 1>(Ap, Cp, Rr);
    st.eA = -Ap[0];
    st.eC = -Cp[0];
    st.eR = Rr[0];
    // slice's first row from lane 0's x_0 = y_0 + p_0 x[st-1] + q_0 e_0
    st.sA = __builtin_fmaf(q[0], st.eA, -pc[0]);
    st.sC = q[0] * st.eC;
    st.sR = __builtin_fmaf(q[0], st.eR, y[0]);
}

// Second half: stage 3 from x[st-1] = xprev, x[ed] = xe and the slice's
// last row se (as the reference's stage 3 takes them).
template <int E>
__device__ __forceinline__ void fast16_stage3(const Fast16State<E>& st, float xprev, float xe, SliceEnd se, int l,
                                              float (&X)[E]) {
    float key = fast_rcp(se.c) * (__builtin_fmaf(-se.a, xprev, se.r) - xe);
    if (se.c == 0.0f) key = 0.0f;
    float e = __builtin_fmaf(-st.eC, key, __builtin_fmaf(-st.eA, xprev, st.eR));
    if (l == 15) e = xe;
    const float eprev = shift_up<16, 1>(e, xprev, l);
#pragma unroll
    for (int k = 0; k < E - 1; k++) X[k] = __builtin_fmaf(st.q[k], e, __builtin_fmaf(st.pc[k], eprev, st.y[k]));
    X[E - 1] = e;
}

template <int E, bool WAVE>
__device__ __forceinline__ void ptpr_solve_fast16_partition(const RawPart<E>& p, bool has_b, int m, int t,
                                                            float* lds, float (&X)[E]) {
    const int l = t & 15;
    const int j = t >> 4;
    Fast16State<E> st;
    fast16_stage1<E>(p, has_b, l, st);
    if constexpr (WAVE) {
        // the system's lanes are t = 0 .. 16 m - 1 of one wavefront
        float xe, xprev;
        SliceEnd se;
        reduced_solve_wave<0, 16>(m, j, (int)(threadIdx.x & 63) - t, st.eA, st.eC, st.eR, st.sA, st.sC, st.sR, xe,
                               xprev, se);
        fast16_stage3<E>(st, xprev, xe, se, l, X);
    } else {
        const float* xr = reduced_solve<0, WAVE>(lds, m, j, l, 16, t, st.eA, st.eC, st.eR, st.sA, st.sC, st.sR);
        const float xprev = (j > 0) ? xr[j - 1] : 0.0f;
        fast16_stage3<E>(st, xprev, xr[j], slice_end(lds, m, j), l, X);
    }
}

// Whole PTPR solve from the raw loads (the division by the diagonal
// included): the two-level form above where it applies, else the
// normalisation followed by the register PCR of ptpr_solve_regs.
template <int E, int L, int MODE, bool WAVE = false>
__device__ __forceinline__ void ptpr_solve_raw(const RawPart<E>& raw, bool has_b, int m, int t, float* lds,
                                               float (&X)[E]) {
    if constexpr (MODE == kFast && L == 16 && E >= 2) {
        ptpr_solve_fast16_partition<E, WAVE>(raw, has_b, m, t, lds, X);
    } else {
        float A[E], C[E], R[E];
        normalise<E, MODE>(raw, has_b, A, C, R);
        ptpr_solve_regs<E, L, MODE, WAVE>(A, C, R, m, t, lds, X);
    }
}

}  // namespace tprh
