// C-ABI harness around the reference's own PTPR class (src/ptpr.hpp:44-110).
// TEST INFRASTRUCTURE ONLY: compiled together with the reference's sources
// (read in place under the reference tree, never copied) into
// oracle/_ref/libtprref.so by oracle/Makefile.  Used by tests and by the
// cpu_baseline leg of bench.py; the product never links it.
#include <cstring>
#include <vector>

#include "ptpr.hpp"

extern "C" {

// One solve through the reference surface: PTPR(n, s),
// set_tridiagonal_system(a, c, rhs) (3-arg, no normalisation), solve(),
// get_ans(x) (src/pm.cpp:182-191).  The caller's arrays are copied first,
// because the reference overwrites them.
int ref_ptpr_solve(const float *a, const float *c, const float *rhs, float *x, int n, int s) {
    std::vector<float> wa(a, a + n), wc(c, c + n), wr(rhs, rhs + n);
    PTPR t(n, s);
    t.set_tridiagonal_system(wa.data(), wc.data(), wr.data());
    int f = t.solve();
    t.get_ans(x);
    return f;
}

// 4-argument setter (normalises in place, src/ptpr.hpp:80-90), applied to
// copies; the normalised arrays are written to na, nc, nr.
void ref_ptpr_normalize(const float *a, const float *diag, const float *c, const float *rhs,
                        float *na, float *nc, float *nr, int n, int s) {
    std::memcpy(na, a, sizeof(float) * n);
    std::memcpy(nc, c, sizeof(float) * n);
    std::memcpy(nr, rhs, sizeof(float) * n);
    std::vector<float> wd(diag, diag + n);
    PTPR t(n, s);
    t.set_tridiagonal_system(na, wd.data(), nc, nr);
}

// Batched solves over nsys systems of n equations (system k at offset k*n).
// mode 0: systems one after another with the reference's OpenMP over slices;
// mode 1: one PTPR instance per thread, OpenMP over systems
// (test/examples.cpp:112-120).  diag may be NULL.
int ref_ptpr_batch(const float *a, const float *diag, const float *c, const float *rhs,
                   float *x, int n, int s, int nsys, int mode) {
    if (mode == 0) {
        PTPR t(n, s);
        std::vector<float> wa(n), wc(n), wr(n), wd(n);
        for (int k = 0; k < nsys; k++) {
            size_t off = (size_t)k * n;
            std::memcpy(wa.data(), a + off, sizeof(float) * n);
            std::memcpy(wc.data(), c + off, sizeof(float) * n);
            std::memcpy(wr.data(), rhs + off, sizeof(float) * n);
            if (diag) {
                std::memcpy(wd.data(), diag + off, sizeof(float) * n);
                t.set_tridiagonal_system(wa.data(), wd.data(), wc.data(), wr.data());
            } else {
                t.set_tridiagonal_system(wa.data(), wc.data(), wr.data());
            }
            t.solve();
            t.get_ans(x + off);
        }
        return 0;
    }
    PTPR proto(n, s);
#pragma omp parallel firstprivate(proto)
    {
        std::vector<float> wa(n), wc(n), wr(n), wd(n);
#pragma omp for schedule(static)
        for (int k = 0; k < nsys; k++) {
            size_t off = (size_t)k * n;
            std::memcpy(wa.data(), a + off, sizeof(float) * n);
            std::memcpy(wc.data(), c + off, sizeof(float) * n);
            std::memcpy(wr.data(), rhs + off, sizeof(float) * n);
            if (diag) {
                std::memcpy(wd.data(), diag + off, sizeof(float) * n);
                proto.set_tridiagonal_system(wa.data(), wd.data(), wc.data(), wr.data());
            } else {
                proto.set_tridiagonal_system(wa.data(), wc.data(), wr.data());
            }
            proto.solve();
            proto.get_ans(x + off);
        }
    }
    return 0;
}
}
