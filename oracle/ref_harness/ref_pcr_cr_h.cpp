// C-ABI harness around the reference's own PCR and CR classes
// (src/pcr.hpp:9-64, src/cr.hpp:24-81), the multi-threaded builds used by
// tpr_pm's `pcr` solver.  TEST INFRASTRUCTURE ONLY (see ref_ptpr_h.cpp).
#include <vector>

#include "cr.hpp"
#include "pcr.hpp"

extern "C" {

int ref_pcr_solve(const float *a, const float *c, const float *rhs, float *x, int n) {
    std::vector<float> wa(a, a + n), wc(c, c + n), wr(rhs, rhs + n);
    PCR p(n);
    p.set_tridiagonal_system(wa.data(), wc.data(), wr.data());
    int f = p.solve();
    p.get_ans(x);
    return f;
}

int ref_cr_solve(const float *a, const float *c, const float *rhs, float *x, int n) {
    std::vector<float> wa(a, a + n), wc(c, c + n), wr(rhs, rhs + n);
    CR s(n);
    s.set_tridiagonal_system(wa.data(), nullptr, wc.data(), wr.data());
    int f = s.solve();
    s.get_ans(x);
    return f;
}
}
