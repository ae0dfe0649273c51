/*
 * tpr_oracle.c -- CPU restatement of the TM-MT/tpr reference solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is linked by
 * the tests, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and by nothing in the product.  The product (libtpr_hip.so)
 * never calls it and has no CPU fallback.
 *
 * Every solver below repeats the reference's floating-point operations in the
 * reference's order, so that a build with -O3 -ffp-contract=off (no FMA, IEEE
 * division, as the reference's Release build on x86-64 without -march) gives
 * results bit-identical to the reference's own classes.  Parity of this file
 * with the reference is pinned in tests/test_oracle.py against
 *   - the reference's compiled sources (oracle/_ref, built by oracle/Makefile),
 *   - golden vectors produced by that build (tests/golden/ref_vectors.npz),
 *   - the reference's known answer test/ans1024.txt,
 *   - closed forms (diagonal systems, FP64 Thomas, dense solves).
 *
 * Function map (reference file:line):
 *   orc_normalize     src/ptpr.hpp:80-90 (4-arg set_tridiagonal_system)
 *   orc_ptpr_solve    src/ptpr.cpp:54-287 (init, solve, stage1, inter, stage2, stage3)
 *   orc_tpr_solve     src/tpr.cpp:54-401  (init, solve, stage1, inter, stage2, st3_replace, stage3)
 *   orc_pcr_solve     src/pcr.cpp:15-101
 *   orc_cr_solve      src/cr.cpp:23-128
 *   orc_norm          src/numerical_stability.cpp:100-115
 *   orc_ptpr_flops    src/ptpr.cpp:104-106; orc_tpr_flops src/tpr.cpp:96-98
 *   orc_gen_*         the seeded synthetic generator (no reference counterpart;
 *                     the reference's RNG is unseeded, src/system.cpp:146-158)
 *   orc_thomas_f64    textbook Thomas algorithm in double (accuracy oracle)
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef float real;
static const real one = 1.0f;
static const real machine_epsilon = FLT_EPSILON; /* src/lib.hpp:13-15 */

static int ilog2i(int v) { /* floor(log2(v)), src/lib.hpp:69-90 */
    int r = -1;
    while (v > 0) { v >>= 1; r++; }
    return r;
}

/* ------------------------------------------------------------------------ */
/* 4-argument setter: divide a, c, rhs by diag in place (src/ptpr.hpp:80-90) */
void orc_normalize(float *a, const float *diag, float *c, float *rhs, int n) {
    for (int i = 0; i < n; i++) {
        a[i] /= diag[i];
        c[i] /= diag[i];
        rhs[i] /= diag[i];
    }
}

/* ------------------------------------------------------------------------ */
/* PCR (src/pcr.cpp:15-89), in place on a, c, rhs; the answer is left in rhs.
 * Work arrays a1, c1, rhs1 have n entries. */
static void pcr_inplace(real *a, real *c, real *rhs, real *a1, real *c1, real *rhs1, int n) {
    int pn = ilog2i(n);
    for (int p = 0; p < pn; p++) {
        int s = 1 << p;
        for (int k = 0; k < s; k++) {
            int kr = k + s;
            real e = one / (one - c[k] * a[kr]);
            a1[k] = e * a[k];
            c1[k] = -e * c[k] * c[kr];
            rhs1[k] = e * (rhs[k] - c[k] * rhs[kr]);
        }
        for (int k = s; k < n - s; k++) {
            int kl = k - s;
            int kr = k + s;
            real ap = a[k];
            real cp = c[k];
            real e = one / (one - ap * c[kl] - cp * a[kr]);
            a1[k] = -e * ap * a[kl];
            c1[k] = -e * cp * c[kr];
            rhs1[k] = e * (rhs[k] - ap * rhs[kl] - cp * rhs[kr]);
        }
        for (int k = n - s; k < n; k++) {
            int kl = k - s;
            real ap = a[k];
            real cp = c[k];
            real e = one / (one - ap * c[kl]);
            a1[k] = -e * ap * a[kl];
            c1[k] = e * cp;
            rhs1[k] = e * (rhs[k] - ap * rhs[kl]);
        }
        for (int k = 0; k < n; k++) {
            a[k] = a1[k];
            c[k] = c1[k];
            rhs[k] = rhs1[k];
        }
    }
}

/* PCR solver on caller data (unit diagonal).  Inputs are not modified. */
int orc_pcr_solve(const float *a_in, const float *c_in, const float *rhs_in, float *x, int n) {
    real *w = (real *)malloc(sizeof(real) * 6 * (size_t)n);
    if (!w) return -1;
    real *a = w, *c = w + n, *rhs = w + 2 * n;
    memcpy(a, a_in, sizeof(real) * n);
    memcpy(c, c_in, sizeof(real) * n);
    memcpy(rhs, rhs_in, sizeof(real) * n);
    pcr_inplace(a, c, rhs, w + 3 * n, w + 4 * n, w + 5 * n, n);
    memcpy(x, rhs, sizeof(real) * n); /* PCR::get_ans, src/pcr.cpp:96-101 */
    free(w);
    return 14 * n * ilog2i(n);
}

/* ------------------------------------------------------------------------ */
/* CR (src/cr.cpp:23-128), in place on a, c, rhs, answer into x.
 * aa, cc, rr are work arrays of n entries. */
static void cr_inplace(real *a, real *c, real *rhs, real *x, real *aa, real *cc, real *rr, int n) {
    if (n <= 1) { /* src/cr.cpp:27-31 */
        x[0] = rhs[0];
        return;
    }
    /* forward reduction, src/cr.cpp:38-82 */
    for (int p = 0; p < ilog2i(n) - 1; p++) {
        int u = 1 << p;
        const int ux = 1 << (p + 1);
        for (int i = ux - 1; i < n - ux; i += ux) {
            int kl = i - u;
            int k = i;
            int kr = i + u;
            real inv_diag_k = one / (one - c[kl] * a[k] - a[kr] * c[k]);
            aa[k] = -inv_diag_k * a[kl] * a[k];
            cc[k] = -inv_diag_k * c[kr] * c[k];
            rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k] - rhs[kr] * c[k]);
        }
        {
            int k = n - 1;
            int kl = k - u;
            real inv_diag_k = one / (one - c[kl] * a[k]);
            aa[k] = -inv_diag_k * a[kl] * a[k];
            cc[k] = inv_diag_k * c[k];
            rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k]);
        }
        for (int i = ux - 1; i < n; i += ux) {
            a[i] = aa[i];
            c[i] = cc[i];
            rhs[i] = rr[i];
        }
    }
    /* backward substitution, src/cr.cpp:88-116 */
    {
        int i = n / 2 - 1;
        int u = n / 2;
        real inv_det = one / (one - c[i] * a[i + u]);
        x[i] = (rhs[i] - c[i] * rhs[i + u]) * inv_det;
        x[i + u] = (rhs[i + u] - rhs[i] * a[i + u]) * inv_det;
    }
    for (int k = ilog2i(n) - 2; k >= 0; k--) {
        int u = 1 << k;
        {
            int i = u - 1;
            x[i] = rhs[i] - c[i] * x[i + u];
        }
        for (int i = u - 1 + 2 * u; i < n - u; i += 2 * u) {
            x[i] = rhs[i] - a[i] * x[i - u] - c[i] * x[i + u];
        }
    }
}

int orc_cr_solve(const float *a_in, const float *c_in, const float *rhs_in, float *x, int n) {
    real *w = (real *)malloc(sizeof(real) * 6 * (size_t)n);
    if (!w) return -1;
    real *a = w, *c = w + n, *rhs = w + 2 * n;
    memcpy(a, a_in, sizeof(real) * n);
    memcpy(c, c_in, sizeof(real) * n);
    memcpy(rhs, rhs_in, sizeof(real) * n);
    cr_inplace(a, c, rhs, x, w + 3 * n, w + 4 * n, w + 5 * n, n);
    free(w);
    return n > 1 ? 17 * n : 0;
}

/* ------------------------------------------------------------------------ */
/* PTPR (src/ptpr.cpp).  Work buffers mirror PTPR::init (src/ptpr.cpp:54-96). */
typedef struct {
    int n, s, m;
    real *a, *c, *rhs;        /* working copies of the caller's arrays */
    real *xbase, *x;          /* x = &xbase[1], x[-1] == 0 */
    real *aa, *cc, *rr;
    real *st2_a, *st2_c, *st2_rhs;
    real *bkup_a, *bkup_c, *bkup_rhs;
    real *pa1, *pc1, *prhs1;  /* PCR work arrays, PCR::init(m) */
} ptpr_ws;

static void ptpr_stage1(ptpr_ws *w, int omp) { /* src/ptpr.cpp:130-213 */
    const int n = w->n, s = w->s;
    real *a = w->a, *c = w->c, *rhs = w->rhs, *aa = w->aa, *cc = w->cc, *rr = w->rr;
#pragma omp parallel for schedule(static) if (omp)
    for (int st = 0; st < n; st += s) {
        for (int p = 1; p <= ilog2i(s); p += 1) {
            int u = 1 << (p - 1);
            int ed = st + s - 1;
            for (int i = st; i < st + u; i++) {
                int k = i;
                int kr = i + u;
                real inv_diag_k = one / (one - a[kr] * c[k]);
                aa[k] = inv_diag_k * a[k];
                cc[k] = -inv_diag_k * c[kr] * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kr] * c[k]);
            }
            for (int i = st + u; i <= ed - u; i++) {
                int kl = i - u;
                int k = i;
                int kr = i + u;
                real inv_diag_k = one / (one - c[kl] * a[k] - a[kr] * c[k]);
                aa[k] = -inv_diag_k * a[kl] * a[k];
                cc[k] = -inv_diag_k * c[kr] * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k] - rhs[kr] * c[k]);
            }
            for (int i = ed - u + 1; i <= ed; i++) {
                int kl = i - u;
                int k = i;
                real inv_diag_k = one / (one - c[kl] * a[k]);
                aa[k] = -inv_diag_k * a[kl] * a[k];
                cc[k] = inv_diag_k * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k]);
            }
            for (int i = st; i <= st + s - 1; i++) {
                a[i] = aa[i];
                c[i] = cc[i];
                rhs[i] = rr[i];
            }
        }
        {
            int ed = st + s - 1;
            real s1;
            if (fabsf(c[ed]) < machine_epsilon) {
                s1 = -1.0f;
            } else {
                s1 = -c[st] / c[ed];
            }
            int dst = st / s;
            w->bkup_a[dst] = a[st];
            w->bkup_c[dst] = c[st];
            w->bkup_rhs[dst] = rhs[st];
            a[st] += s1 * a[ed];
            c[st] = s1;
            rhs[st] += s1 * rhs[ed];
        }
    }
}

static void ptpr_inter(ptpr_ws *w) { /* src/ptpr.cpp:218-241 */
    const int n = w->n, s = w->s, m = w->m;
    for (int i = s - 1; i < n - 1; i += s) {
        int k = i;
        int kr = i + 1;
        real ak = w->a[k], akr = w->a[kr];
        real ck = w->c[k], ckr = w->c[kr];
        real rhsk = w->rhs[k], rhskr = w->rhs[kr];
        real inv_diag_k = one / (one - akr * ck);
        int dst = i / s;
        w->st2_a[dst] = inv_diag_k * ak;
        w->st2_c[dst] = -inv_diag_k * ckr * ck;
        w->st2_rhs[dst] = inv_diag_k * (rhsk - rhskr * ck);
    }
    w->st2_a[m - 1] = w->a[n - 1];
    w->st2_c[m - 1] = w->c[n - 1];
    w->st2_rhs[m - 1] = w->rhs[n - 1];
}

static void ptpr_stage2(ptpr_ws *w) { /* src/ptpr.cpp:246-258 */
    pcr_inplace(w->st2_a, w->st2_c, w->st2_rhs, w->pa1, w->pc1, w->prhs1, w->m);
    for (int i = w->s - 1; i < w->n; i += w->s) w->x[i] = w->st2_rhs[i / w->s];
}

static void ptpr_stage3(ptpr_ws *w, int omp) { /* src/ptpr.cpp:263-287 */
    const int n = w->n, s = w->s;
    real *a = w->a, *c = w->c, *rhs = w->rhs, *x = w->x;
#pragma omp parallel for schedule(static) if (omp)
    for (int st = 0; st < n; st += s) {
        int src = st / s;
        a[st] = w->bkup_a[src];
        c[st] = w->bkup_c[src];
        rhs[st] = w->bkup_rhs[src];
        int ed = st + s - 1;
        real key = one / c[ed] * (rhs[ed] - a[ed] * x[st - 1] - x[ed]);
        if (c[ed] == 0.0f) key = 0.0f;
        for (int i = st; i < ed; i++) x[i] = rhs[i] - a[i] * x[st - 1] - c[i] * key;
    }
}

static int ptpr_ws_init(ptpr_ws *w, int n, int s) {
    memset(w, 0, sizeof(*w));
    w->n = n; w->s = s; w->m = n / s;
    size_t need = (size_t)3 * n + (n + 1) + (size_t)3 * n + (size_t)9 * w->m;
    real *buf = (real *)calloc(need, sizeof(real));
    if (!buf) return -1;
    real *p = buf;
    w->a = p; p += n; w->c = p; p += n; w->rhs = p; p += n;
    w->xbase = p; p += n + 1; w->x = w->xbase + 1;
    w->aa = p; p += n; w->cc = p; p += n; w->rr = p; p += n;
    w->st2_a = p; p += w->m; w->st2_c = p; p += w->m; w->st2_rhs = p; p += w->m;
    w->bkup_a = p; p += w->m; w->bkup_c = p; p += w->m; w->bkup_rhs = p; p += w->m;
    w->pa1 = p; p += w->m; w->pc1 = p; p += w->m; w->prhs1 = p; p += w->m;
    return 0;
}

int orc_ptpr_flops(int n, int s) { /* src/ptpr.cpp:104-106 */
    int m = n / s;
    int fp_st1 = m * (14 * s * ilog2i(s)) + 14 * m;
    int fp_st2 = 9 * (m - 1) + (14 * m * ilog2i(m));
    int fp_st3 = m * 4 * (s - 1);
    return fp_st1 + fp_st2 + fp_st3;
}

static int ptpr_run(ptpr_ws *w, const float *a, const float *c, const float *rhs, float *x, int omp) {
    memcpy(w->a, a, sizeof(real) * w->n);
    memcpy(w->c, c, sizeof(real) * w->n);
    memcpy(w->rhs, rhs, sizeof(real) * w->n);
    ptpr_stage1(w, omp);
    ptpr_inter(w);
    ptpr_stage2(w);
    ptpr_stage3(w, omp);
    memcpy(x, w->x, sizeof(real) * w->n); /* PTPR::get_ans, src/ptpr.cpp:362-367 */
    return orc_ptpr_flops(w->n, w->s);
}

/* One PTPR solve with the unit-diagonal system (a, c, rhs); x receives the
 * answer.  Inputs are left untouched (the reference overwrites them). */
int orc_ptpr_solve(const float *a, const float *c, const float *rhs, float *x, int n, int s) {
    ptpr_ws w;
    if (ptpr_ws_init(&w, n, s)) return -1;
    int f = ptpr_run(&w, a, c, rhs, x, 1);
    free(w.a);
    return f;
}

/* ------------------------------------------------------------------------ */
/* TPR (src/tpr.cpp). */
typedef struct {
    int n, s, m;
    real *a, *c, *rhs, *x;
    real *aa, *cc, *rr;
    real *st2_a, *st2_c, *st2_rhs;
    real *bkup_a, *bkup_c, *bkup_rhs;
    real *cr_aa, *cr_cc, *cr_rr, *cr_x; /* CR::init(m) */
} tpr_ws;

static void tpr_stage1(tpr_ws *w, int omp) { /* src/tpr.cpp:123-230 */
    const int n = w->n, s = w->s;
    real *a = w->a, *c = w->c, *rhs = w->rhs, *aa = w->aa, *cc = w->cc, *rr = w->rr;
#pragma omp parallel for schedule(static) if (omp)
    for (int st = 0; st < n; st += s) {
        for (int i = st; i <= st + s - 1; i += 2) {
            w->bkup_a[i] = a[i];
            w->bkup_c[i] = c[i];
            w->bkup_rhs[i] = rhs[i];
        }
        for (int p = 1; p <= ilog2i(s); p += 1) {
            int u = 1 << (p - 1);
            int p2k = 1 << p;
            int ed = st + s - 1;
            {
                int k = st;
                int kr = st + u;
                real inv_diag_k = one / (one - a[kr] * c[k]);
                aa[k] = inv_diag_k * a[k];
                cc[k] = -inv_diag_k * c[kr] * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kr] * c[k]);
            }
            for (int i = st + p2k; i <= ed - u; i += p2k) {
                int kl = i - u;
                int k = i;
                int kr = i + u;
                real inv_diag_k = one / (one - c[kl] * a[k] - a[kr] * c[k]);
                aa[k] = -inv_diag_k * a[kl] * a[k];
                cc[k] = -inv_diag_k * c[kr] * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k] - rhs[kr] * c[k]);
            }
            for (int i = st + p2k - 1; i <= ed - u; i += p2k) {
                int kl = i - u;
                int k = i;
                int kr = i + u;
                real inv_diag_k = one / (one - c[kl] * a[k] - a[kr] * c[k]);
                aa[k] = -inv_diag_k * a[kl] * a[k];
                cc[k] = -inv_diag_k * c[kr] * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k] - rhs[kr] * c[k]);
            }
            {
                int kl = ed - u;
                int k = ed;
                real inv_diag_k = one / (one - c[kl] * a[k]);
                aa[k] = -inv_diag_k * a[kl] * a[k];
                cc[k] = inv_diag_k * c[k];
                rr[k] = inv_diag_k * (rhs[k] - rhs[kl] * a[k]);
            }
            for (int i = st; i <= ed; i += p2k) {
                a[i] = aa[i];
                c[i] = cc[i];
                rhs[i] = rr[i];
            }
            for (int i = st + p2k - 1; i <= ed; i += p2k) {
                a[i] = aa[i];
                c[i] = cc[i];
                rhs[i] = rr[i];
            }
        }
        for (int i = st + 1; i <= st + s - 1; i += 2) {
            w->bkup_a[i] = a[i];
            w->bkup_c[i] = c[i];
            w->bkup_rhs[i] = rhs[i];
        }
        {
            int ed = st + s - 1;
            real s1;
            if (fabsf(c[ed]) < machine_epsilon) {
                s1 = -1.0f;
            } else {
                s1 = -c[st] / c[ed];
            }
            a[st] += s1 * a[ed];
            c[st] = s1;
            rhs[st] += s1 * rhs[ed];
        }
    }
}

static void tpr_inter(tpr_ws *w) { /* src/tpr.cpp:235-258 */
    const int n = w->n, s = w->s, m = w->m;
    for (int i = s - 1; i < n - 1; i += s) {
        int k = i;
        int kr = i + 1;
        real ak = w->a[k], akr = w->a[kr];
        real ck = w->c[k], ckr = w->c[kr];
        real rhsk = w->rhs[k], rhskr = w->rhs[kr];
        real inv_diag_k = one / (one - akr * ck);
        int dst = i / s;
        w->st2_a[dst] = inv_diag_k * ak;
        w->st2_c[dst] = -inv_diag_k * ckr * ck;
        w->st2_rhs[dst] = inv_diag_k * (rhsk - rhskr * ck);
    }
    w->st2_a[m - 1] = w->a[n - 1];
    w->st2_c[m - 1] = w->c[n - 1];
    w->st2_rhs[m - 1] = w->rhs[n - 1];
}

static void tpr_stage2(tpr_ws *w) { /* src/tpr.cpp:263-275 */
    cr_inplace(w->st2_a, w->st2_c, w->st2_rhs, w->cr_x, w->cr_aa, w->cr_cc, w->cr_rr, w->m);
    memcpy(w->st2_rhs, w->cr_x, sizeof(real) * w->m); /* CR::get_ans into st2_rhs */
    for (int i = 0; i < w->m; i++) w->x[(i + 1) * w->s - 1] = w->st2_rhs[i];
}

static void tpr_stage3(tpr_ws *w, int omp) { /* src/tpr.cpp:381-387 then :277-308 */
    const int n = w->n, s = w->s;
    for (int i = 0; i < n; i++) { /* st3_replace */
        w->a[i] = w->bkup_a[i];
        w->c[i] = w->bkup_c[i];
        w->rhs[i] = w->bkup_rhs[i];
    }
    real *a = w->a, *c = w->c, *rhs = w->rhs, *x = w->x;
#pragma omp parallel for if (omp)
    for (int st = 0; st < n; st += s) {
        for (int p = ilog2i(s) - 1; p >= 0; p--) {
            int ed = st + s - 1;
            int u = 1 << p;
            {
                int i = st + u - 1;
                real x_u;
                if (i - u < 0) {
                    x_u = 0.0f;
                } else {
                    x_u = x[i - u];
                }
                x[i] = rhs[i] - a[i] * x_u - c[i] * x[i + u];
            }
            for (int i = st + u - 1 + 2 * u; i <= ed; i += 2 * u) {
                x[i] = rhs[i] - a[i] * x[i - u] - c[i] * x[i + u];
            }
        }
    }
}

int orc_tpr_flops(int n, int s) { /* src/tpr.cpp:96-98 */
    int m = n / s;
    return (28 * n - 14 * m) + (9 * (m - 1) + 17 * m) + m * 4 * (s - 1);
}

static int tpr_ws_init(tpr_ws *w, int n, int s) {
    memset(w, 0, sizeof(*w));
    w->n = n; w->s = s; w->m = n / s;
    size_t need = (size_t)10 * n + (size_t)7 * w->m;
    real *buf = (real *)calloc(need, sizeof(real));
    if (!buf) return -1;
    real *p = buf;
    w->a = p; p += n; w->c = p; p += n; w->rhs = p; p += n; w->x = p; p += n;
    w->aa = p; p += n; w->cc = p; p += n; w->rr = p; p += n;
    w->bkup_a = p; p += n; w->bkup_c = p; p += n; w->bkup_rhs = p; p += n;
    w->st2_a = p; p += w->m; w->st2_c = p; p += w->m; w->st2_rhs = p; p += w->m;
    w->cr_aa = p; p += w->m; w->cr_cc = p; p += w->m; w->cr_rr = p; p += w->m; w->cr_x = p; p += w->m;
    return 0;
}

static int tpr_run(tpr_ws *w, const float *a, const float *c, const float *rhs, float *x, int omp) {
    memcpy(w->a, a, sizeof(real) * w->n);
    memcpy(w->c, c, sizeof(real) * w->n);
    memcpy(w->rhs, rhs, sizeof(real) * w->n);
    tpr_stage1(w, omp);
    tpr_inter(w);
    tpr_stage2(w);
    tpr_stage3(w, omp);
    memcpy(x, w->x, sizeof(real) * w->n); /* TPR::get_ans, src/tpr.cpp:396-401 */
    return orc_tpr_flops(w->n, w->s);
}

int orc_tpr_solve(const float *a, const float *c, const float *rhs, float *x, int n, int s) {
    tpr_ws w;
    if (tpr_ws_init(&w, n, s)) return -1;
    int f = tpr_run(&w, a, c, rhs, x, 1);
    free(w.a);
    return f;
}

/* ------------------------------------------------------------------------ */
/* Batched solves over nsys back-to-back systems of n equations (system k at
 * offset k*n).  mode 0 = "reference shape": systems one after another, OpenMP
 * over slices inside each solve (src/ptpr.cpp:131,264); mode 1 = "throughput
 * shape": one solver per thread, OpenMP over systems (the pattern of
 * test/examples.cpp:112-120).  diag may be NULL (unit diagonal); when given,
 * each system is normalised first, as the 4-argument setter does.
 * method: 0 = PTPR, 1 = TPR. */
int orc_solve_batch(int method, const float *a, const float *diag, const float *c,
                    const float *rhs, float *x, int n, int s, int nsys, int mode) {
    int err = 0;
    if (mode == 0) {
        ptpr_ws pw; tpr_ws tw;
        if (method == 0 ? ptpr_ws_init(&pw, n, s) : tpr_ws_init(&tw, n, s)) return -1;
        float *na = NULL, *nc = NULL, *nr = NULL;
        if (diag) {
            na = (float *)malloc(sizeof(float) * 3 * (size_t)n);
            nc = na + n; nr = nc + n;
        }
        for (int k = 0; k < nsys; k++) {
            size_t off = (size_t)k * n;
            const float *pa = a + off, *pc = c + off, *pr = rhs + off;
            if (diag) {
                memcpy(na, pa, sizeof(float) * n);
                memcpy(nc, pc, sizeof(float) * n);
                memcpy(nr, pr, sizeof(float) * n);
                orc_normalize(na, diag + off, nc, nr, n);
                pa = na; pc = nc; pr = nr;
            }
            if (method == 0) ptpr_run(&pw, pa, pc, pr, x + off, 1);
            else tpr_run(&tw, pa, pc, pr, x + off, 1);
        }
        free(na);
        if (method == 0) free(pw.a); else free(tw.a);
        return 0;
    }
#pragma omp parallel reduction(| : err)
    {
        ptpr_ws pw; tpr_ws tw;
        int ok = (method == 0 ? ptpr_ws_init(&pw, n, s) : tpr_ws_init(&tw, n, s)) == 0;
        float *na = (float *)malloc(sizeof(float) * 3 * (size_t)n);
        if (!ok || !na) {
            err |= 1;
        } else {
            float *nc = na + n, *nr = nc + n;
#pragma omp for schedule(static)
            for (int k = 0; k < nsys; k++) {
                size_t off = (size_t)k * n;
                const float *pa = a + off, *pc = c + off, *pr = rhs + off;
                if (diag) {
                    memcpy(na, pa, sizeof(float) * n);
                    memcpy(nc, pc, sizeof(float) * n);
                    memcpy(nr, pr, sizeof(float) * n);
                    orc_normalize(na, diag + off, nc, nr, n);
                    pa = na; pc = nc; pr = nr;
                }
                if (method == 0) ptpr_run(&pw, pa, pc, pr, x + off, 0);
                else tpr_run(&tw, pa, pc, pr, x + off, 0);
            }
        }
        free(na);
        if (ok) { if (method == 0) free(pw.a); else free(tw.a); }
    }
    return err ? -1 : 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Relative L2 error, src/numerical_stability.cpp:100-115: the difference and
 * its square in `real`, the sums in double. */
double orc_norm(const float *solution, const float *xtruth, int n) {
    double sum1 = 0.0, sum2 = 0.0;
    for (int i = 0; i < n; i++) {
        real d = solution[i] - xtruth[i];
        double dsq = (double)(d * d);
        sum1 += dsq;
        sum2 += xtruth[i] * xtruth[i];
    }
    return sqrt(sum1) / sqrt(sum2);
}

/* ------------------------------------------------------------------------ */
/* FP64 Thomas algorithm (no pivoting) on a general tridiagonal system
 * a[i] x[i-1] + b[i] x[i] + c[i] x[i+1] = d[i], inputs in float, all
 * arithmetic and the answer in double.  w: work array of n doubles. */
void orc_thomas_f64(const float *a, const float *b, const float *c, const float *d,
                    double *x, int n) {
    double *cp = (double *)malloc(sizeof(double) * (size_t)n);
    double beta = (double)b[0];
    cp[0] = (double)c[0] / beta;
    x[0] = (double)d[0] / beta;
    for (int i = 1; i < n; i++) {
        beta = (double)b[i] - (double)a[i] * cp[i - 1];
        cp[i] = (double)c[i] / beta;
        x[i] = ((double)d[i] - (double)a[i] * x[i - 1]) / beta;
    }
    for (int i = n - 2; i >= 0; i--) x[i] -= cp[i] * x[i + 1];
    free(cp);
}

/* Batched FP64 Thomas over nsys systems (OpenMP over systems). */
void orc_thomas_f64_batch(const float *a, const float *b, const float *c, const float *d,
                          double *x, int n, int nsys) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nsys; k++) {
        size_t off = (size_t)k * n;
        orc_thomas_f64(a + off, b + off, c + off, d + off, x + off, n);
    }
}

/* FP64 residual r = d - A x of a float system at a double x. */
void orc_residual_f64(const float *a, const float *b, const float *c, const float *d,
                      const double *x, double *r, int n) {
    for (int i = 0; i < n; i++) {
        double acc = (double)d[i] - (double)b[i] * x[i];
        if (i > 0) acc -= (double)a[i] * x[i - 1];
        if (i < n - 1) acc -= (double)c[i] * x[i + 1];
        r[i] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* Seeded counter-based generator.  The same bits are produced by the HIP
 * generator kernel; every value depends only on (seed, system id, stream,
 * index), never on batch composition.
 *
 * draw(seed, sys, stream, i) = mix64(key(seed, sys) + G * (8*i + stream + 1))
 * key(seed, sys)             = mix64(seed + G * (sys + 1)),  G = 0x9E3779B97F4A7C15
 * U(-1,1) from the top 24 bits k: (k - 2^23) * 2^-23 (exact in fp32);
 * U[0,1)  from the top 24 bits k: k * 2^-24.                                */
static inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ULL;
    z ^= z >> 27; z *= 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z;
}
#define GOLDEN 0x9E3779B97F4A7C15ULL
static inline uint64_t sys_key(uint64_t seed, uint64_t sys) { return mix64(seed + GOLDEN * (sys + 1)); }
static inline uint32_t draw24(uint64_t key, uint64_t stream, uint64_t i) {
    return (uint32_t)(mix64(key + GOLDEN * (8 * i + stream + 1)) >> 40);
}
static inline float u_pm1(uint32_t k) { return (float)((int32_t)k - 8388608) * (1.0f / 8388608.0f); }
static inline float u_01(uint32_t k) { return (float)k * (1.0f / 16777216.0f); }

/* Streams */
enum { ST_A = 0, ST_C = 1, ST_B = 2, ST_XT = 3, ST_RHS = 4 };

/* kind: 0 = strictly dominant (b = |a|+|c| + (1 + U[0,1)));
 *       1 = weakly dominant   (b = |a|+|c| + (1e-4 + U[0,1) * 0.0099));
 *       2 = ExampleFixedInput (a = c = -1/6, b = 1, rhs = i + 1; src/system.cpp:118-130);
 *       3 = ExampleRandomRHSInput (a = c = -1/6, b = 1, rhs ~ U(-1,1); src/system.cpp:146-158).
 * For kinds 0 and 1, xt ~ U(-1,1) and d = A*xt accumulated in double, rounded
 * to float.  For kinds 2 and 3, xt is left untouched (may be NULL). */
void orc_gen_system(uint64_t seed, uint64_t sys, int n, int kind,
                    float *a, float *b, float *c, float *d, float *xt) {
    uint64_t key = sys_key(seed, sys);
    if (kind == 2 || kind == 3) {
        for (int i = 0; i < n; i++) {
            a[i] = (float)(-1.0 / 6.0);
            c[i] = (float)(-1.0 / 6.0);
            b[i] = 1.0f;
            d[i] = kind == 2 ? (float)(1.0 * (i + 1)) : u_pm1(draw24(key, ST_RHS, (uint64_t)i));
        }
        a[0] = 0.0f;
        c[n - 1] = 0.0f;
        return;
    }
    for (int i = 0; i < n; i++) {
        float ai = u_pm1(draw24(key, ST_A, (uint64_t)i));
        float ci = u_pm1(draw24(key, ST_C, (uint64_t)i));
        if (i == 0) ai = 0.0f;
        if (i == n - 1) ci = 0.0f;
        float u = u_01(draw24(key, ST_B, (uint64_t)i));
        float extra = kind == 0 ? (1.0f + u) : (1e-4f + u * 0.0099f);
        a[i] = ai;
        c[i] = ci;
        b[i] = (fabsf(ai) + fabsf(ci)) + extra;
        xt[i] = u_pm1(draw24(key, ST_XT, (uint64_t)i));
    }
    for (int i = 0; i < n; i++) {
        double acc = (double)b[i] * (double)xt[i];
        if (i > 0) acc = (double)a[i] * (double)xt[i - 1] + acc;
        if (i < n - 1) acc = acc + (double)c[i] * (double)xt[i + 1];
        d[i] = (float)acc;
    }
}

/* nsys systems with ids sys0 .. sys0+nsys-1, back to back. */
void orc_gen_batch(uint64_t seed, uint64_t sys0, int nsys, int n, int kind,
                   float *a, float *b, float *c, float *d, float *xt) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nsys; k++) {
        size_t off = (size_t)k * n;
        orc_gen_system(seed, sys0 + k, n, kind, a + off, b + off, c + off, d + off,
                       xt ? xt + off : NULL);
    }
}

/* Right-hand side number j of system `sys` (multi-RHS requests): d ~ U(-1,1). */
void orc_gen_rhs(uint64_t seed, uint64_t sys, int j, int n, float *d) {
    uint64_t key = sys_key(seed, sys);
    for (int i = 0; i < n; i++)
        d[i] = u_pm1(draw24(key, 5 + 0, (uint64_t)j * (uint64_t)n + (uint64_t)i));
}
