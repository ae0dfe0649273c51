// C-ABI harness around the reference's own TPR class (src/tpr.hpp:44-112).
// TEST INFRASTRUCTURE ONLY (see ref_ptpr_h.cpp).
#include <cstring>
#include <vector>

#include "tpr.hpp"

extern "C" {

int ref_tpr_solve(const float *a, const float *c, const float *rhs, float *x, int n, int s) {
    std::vector<float> wa(a, a + n), wc(c, c + n), wr(rhs, rhs + n);
    TPR t(n, s);
    t.set_tridiagonal_system(wa.data(), wc.data(), wr.data());
    int f = t.solve();
    t.get_ans(x);
    return f;
}

int ref_tpr_batch(const float *a, const float *diag, const float *c, const float *rhs,
                  float *x, int n, int s, int nsys, int mode) {
    if (mode == 0) {
        TPR t(n, s);
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
    TPR proto(n, s);
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
