// known_answer.cpp -- the reference's GPU example tests (test/cuexamples.cpp:
// 9-83, CR / PCR / TPR / PTPR on ExampleFixedInput(1024) for s = 64 .. 1024)
// and its class tests (test/examples.cpp:36-102), restated against the
// MI355X engine in C++: the CUDA entry points through tpr::hip::*_cu, the
// classes through tpr::hip::PTPR / TPR / PCR / CR.  Same known answer
// (test/ans1024.txt, read here from its committed copy ans1024.npy) and the
// same acceptance: |x_i - ans_i| <= 1e-2 for every i and
// max_i (x_i - ans_i)^2 <= 1e-3.
//
//   known_answer <path to ans1024.npy>      (exit status 0 when every case passes)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "tpr/hip_solvers.hpp"

namespace {

constexpr int kN = 1024;
int g_failures = 0;

// A one-dimensional little-endian float64 .npy file.
bool read_npy_f64(const char* path, std::vector<double>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    char magic[6];
    f.read(magic, 6);
    if (std::memcmp(magic, "\x93NUMPY", 6) != 0) return false;
    unsigned char ver[2];
    f.read(reinterpret_cast<char*>(ver), 2);
    unsigned int hlen = 0;
    if (ver[0] == 1) {
        unsigned char h[2];
        f.read(reinterpret_cast<char*>(h), 2);
        hlen = h[0] | (h[1] << 8);
    } else {
        unsigned char h[4];
        f.read(reinterpret_cast<char*>(h), 4);
        hlen = h[0] | (h[1] << 8) | (h[2] << 16) | ((unsigned)h[3] << 24);
    }
    std::string header(hlen, ' ');
    f.read(&header[0], hlen);
    if (header.find("'<f8'") == std::string::npos || header.find("True") != std::string::npos) return false;
    out.resize(kN);
    f.read(reinterpret_cast<char*>(out.data()), sizeof(double) * kN);
    return (bool)f;
}

// ExampleFixedInput (src/system.cpp:118-130): a = c = -1/6, diag = 1, rhs = i + 1.
struct FixedInput {
    std::vector<float> a, diag, c, rhs;
    FixedInput() : a(kN), diag(kN), c(kN), rhs(kN) { assign(); }
    void assign() {
        for (int i = 0; i < kN; i++) {
            a[i] = (float)(-1.0 / 6.0);
            c[i] = (float)(-1.0 / 6.0);
            diag[i] = 1.0f;
            rhs[i] = (float)(1.0 * (i + 1));
        }
        a[0] = 0.0f;
        c[kN - 1] = 0.0f;
    }
};

void check(const char* name, int s, const std::vector<double>& ans, const std::vector<float>& x) {
    float maxsq = 0.0f;
    int worst = -1, far = 0;
    for (int i = 0; i < kN; i++) {
        const float d = std::fabs((float)ans[i] - x[i]);
        if (!(d <= 1e-2f)) far++;
        const float sq = d * d;
        if (sq > maxsq || worst < 0) {
            maxsq = sq;
            worst = i;
        }
    }
    const bool ok = far == 0 && maxsq <= 1e-3f;
    if (!ok) g_failures++;
    std::printf("%-5s %-10s s=%-5d (|x - ans| > 1e-2 at %d entries, max sq %.3g at %d)\n", ok ? "ok" : "FAIL",
                name, s, far, maxsq, worst);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: known_answer <ans1024.npy>\n");
        return 2;
    }
    std::vector<double> ans;
    if (!read_npy_f64(argv[1], ans)) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    try {
        FixedInput in;
        std::vector<float> x(kN);
        // test/cuexamples.cpp: the CUDA entry points (unit diagonal, host arrays)
        tpr::hip::cr_cu(in.a.data(), in.c.data(), in.rhs.data(), x.data(), kN);
        check("cr_cu", 0, ans, x);
        tpr::hip::pcr_cu(in.a.data(), in.c.data(), in.rhs.data(), x.data(), kN);
        check("pcr_cu", 0, ans, x);
        for (int s = 64; s <= kN; s *= 2) {
            in.assign();
            tpr::hip::tpr_cu(in.a.data(), in.c.data(), in.rhs.data(), x.data(), kN, s);
            check("tpr_cu", s, ans, x);
            in.assign();
            tpr::hip::ptpr_cu(in.a.data(), in.c.data(), in.rhs.data(), x.data(), kN, s);
            check("ptpr_cu", s, ans, x);
        }
        // test/examples.cpp: the solver classes, 6-argument constructors
        {
            in.assign();
            tpr::hip::CR cr(in.a.data(), in.diag.data(), in.c.data(), in.rhs.data(), kN);
            cr.solve();
            cr.get_ans(x.data());
            check("CR", 0, ans, x);
            in.assign();
            tpr::hip::PCR pcr(in.a.data(), in.diag.data(), in.c.data(), in.rhs.data(), kN);
            pcr.solve();
            pcr.get_ans(x.data());
            check("PCR", 0, ans, x);
        }
        for (int s = 4; s <= kN; s *= 2) {
            in.assign();
            tpr::hip::TPR t(in.a.data(), in.diag.data(), in.c.data(), in.rhs.data(), kN, s);
            t.solve();
            t.get_ans(x.data());
            check("TPR", s, ans, x);
            in.assign();
            tpr::hip::PTPR p(in.a.data(), in.diag.data(), in.c.data(), in.rhs.data(), kN, s);
            p.solve();
            p.get_ans(x.data());
            check("PTPR", s, ans, x);
        }
    } catch (const tpr::hip::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("%d failure(s)\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
