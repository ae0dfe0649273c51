// A caller written against the reference's own declarations (src/ptpr.cuh,
// src/tpr.cuh), as test/cuexamples.cpp:49-83 calls them: it compiles and
// links unchanged against the shims above and libtpr_hip.so.  Run on a GPU
// it solves ExampleFixedInput(1024) with every entry point and checks the
// known answer's pattern x_i ~ 1.5 (i + 1) away from the ends.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ptpr.cuh"
#include "tpr.cuh"

int main() {
    const int n = 1024;
    std::vector<float> a(n, -1.0f / 6.0f), c(n, -1.0f / 6.0f), rhs(n), x(n);
    a[0] = 0.0f;
    c[n - 1] = 0.0f;
    for (int i = 0; i < n; i++) rhs[i] = (float)(i + 1);
    int bad = 0;
    for (int s = 64; s <= n; s *= 2) {
        PTPR_CU::ptpr_cu(a.data(), c.data(), rhs.data(), x.data(), n, s);
        bad += std::fabs(x[n / 2] - 1.5f * (n / 2 + 1)) > 1e-2f;
        TPR_CU::tpr_cu(a.data(), c.data(), rhs.data(), x.data(), n, s);
        bad += std::fabs(x[n / 2] - 1.5f * (n / 2 + 1)) > 1e-2f;
    }
    PTPR_CU::pcr_cu(a.data(), c.data(), rhs.data(), x.data(), n);
    bad += std::fabs(x[n / 2] - 1.5f * (n / 2 + 1)) > 1e-2f;
    TPR_CU::cr_cu(a.data(), c.data(), rhs.data(), x.data(), n);
    bad += std::fabs(x[n / 2] - 1.5f * (n / 2 + 1)) > 1e-2f;
    std::printf("%s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
