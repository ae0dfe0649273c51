// tpr_nstab -- accuracy report over dense-CSV tridiagonal inputs, with the
// reference nstab driver's protocol and output (src/numerical_stability.cpp:
// 143-195), on the MI355X solvers.
//
//   tpr_nstab [INPUT_DIR] [--fast]
//
// INPUT_DIR holds eq*.txt (n = 512 rows of n + 1 comma-separated values:
// the matrix row, then the right-hand side; read in double, cast to fp32 as
// src/system.cpp:54-101 does) and xt.txt (the truth, one value per line).
// For every file (in glob order) it prints, as CSV `file,solver,N,S,L2`:
// PTPR and TPR at s = 16, 32, ..., n, PCR (S = 0), and the rocSPARSE
// comparator REFERENCE_ROCSPARSE (S = 0) when libtpr_rocsparse.so loads.
// L2 = ||x - xt||_2 / ||xt||_2 with the reference's mixed-precision sums
// (src/numerical_stability.cpp:100-115).  Every solver reads a freshly
// loaded copy of the system (the reference reloads before each solver
// because its setters divide the arrays in place).
// Arithmetic: exact (bit-identical to the reference's CPU classes) unless
// --fast.
#include <dlfcn.h>
#include <glob.h>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tpr/hip_solvers.hpp"

namespace {

struct System {
    std::vector<float> a, diag, c, rhs;
};

bool read_system(const std::string& path, int n, System& s) {
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return false;
    std::vector<double> a(n), d(n), c(n), r(n);
    double v;
    bool ok = true;
    for (int i = 0; i < n && ok; i++) {
        for (int j = 0; j < n && ok; j++) {
            ok = std::fscanf(fp, "%lf,", &v) == 1;
            if (j == i - 1) a[i] = v;
            if (j == i) d[i] = v;
            if (j == i + 1) c[i] = v;
        }
        ok = ok && std::fscanf(fp, "%lf", &r[i]) == 1;
    }
    std::fclose(fp);
    if (!ok) return false;
    a[0] = 0.0;
    c[n - 1] = 0.0;
    s.a.assign(a.begin(), a.end());
    s.diag.assign(d.begin(), d.end());
    s.c.assign(c.begin(), c.end());
    s.rhs.assign(r.begin(), r.end());
    return true;
}

bool read_truth(const std::string& path, int n, std::vector<float>& xt) {
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp) return false;
    xt.resize(n);
    double v;
    for (int i = 0; i < n; i++) {
        if (std::fscanf(fp, "%lf", &v) != 1) {
            std::fclose(fp);
            return false;
        }
        xt[i] = (float)v;
    }
    std::fclose(fp);
    return true;
}

double rel_l2(const std::vector<float>& x, const std::vector<float>& xt) {
    const size_t n = x.size();
    std::vector<double> dsq(n);
    for (size_t i = 0; i < n; i++) {
        const float df = x[i] - xt[i];
        dsq[i] = df * df;  // fp32 square, as the reference's real arithmetic
    }
    double s1 = 0.0, s2 = 0.0;
    for (size_t i = 0; i < n; i++) {
        s1 += dsq[i];
        s2 += xt[i] * xt[i];
    }
    return std::sqrt(s1) / std::sqrt(s2);
}

std::string base_name(const std::string& p) {
    const size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

template <class Solver>
double eval(Solver& solver, const System& sys0, const std::vector<float>& xt, bool fast) {
    System sys = sys0;
    if (fast) solver.set_arith(TPRH_ARITH_FAST);
    solver.set_tridiagonal_system(sys.a.data(), sys.diag.data(), sys.c.data(), sys.rhs.data());
    solver.solve();
    std::vector<float> x(sys.a.size());
    solver.get_ans(x.data());
    return rel_l2(x, xt);
}

// The rocSPARSE slot, looked up at run time so the report runs without it.
using VendorFn = int (*)(const float*, const float*, const float*, const float*, float*, int, int, void*);

VendorFn load_vendor() {
    Dl_info info;
    std::string dir = ".";
    if (dladdr(reinterpret_cast<void*>(&tprh_version), &info) && info.dli_fname) {
        std::string f = info.dli_fname;
        const size_t k = f.find_last_of('/');
        if (k != std::string::npos) dir = f.substr(0, k);
    }
    void* h = dlopen((dir + "/libtpr_rocsparse.so").c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) return nullptr;
    return reinterpret_cast<VendorFn>(dlsym(h, "tprh_vendor_gtsv_batch"));
}

double eval_vendor(VendorFn f, const System& s, const std::vector<float>& xt) {
    const int n = (int)s.a.size();
    const size_t nb = sizeof(float) * n;
    float* dev = nullptr;
    if (hipMalloc(&dev, 5 * nb) != hipSuccess) return NAN;
    const float* src[4] = {s.a.data(), s.diag.data(), s.c.data(), s.rhs.data()};
    for (int k = 0; k < 4; k++) (void)hipMemcpy(dev + (size_t)k * n, src[k], nb, hipMemcpyHostToDevice);
    std::vector<float> x(n);
    int st = f(dev, dev + n, dev + 2 * (size_t)n, dev + 3 * (size_t)n, dev + 4 * (size_t)n, n, 1, nullptr);
    if (st == 0) st = hipMemcpy(x.data(), dev + 4 * (size_t)n, nb, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
    (void)hipFree(dev);
    return st == 0 ? rel_l2(x, xt) : NAN;
}

}  // namespace

int main(int argc, char** argv) {
    std::string dir = ".";
    bool fast = false;
    for (int i = 1; i < argc; i++) {
        if (std::strcmp(argv[i], "--fast") == 0) fast = true;
        else dir = argv[i];
    }
    const int n = 512;
    std::vector<float> xt;
    if (!read_truth(dir + "/xt.txt", n, xt)) {
        std::fprintf(stderr, "Failed to read %s/xt.txt\n", dir.c_str());
        return EXIT_FAILURE;
    }
    glob_t g;
    std::memset(&g, 0, sizeof(g));
    if (glob((dir + "/eq*.txt").c_str(), GLOB_TILDE, nullptr, &g) != 0) {
        std::fprintf(stderr, "no eq*.txt under %s\n", dir.c_str());
        globfree(&g);
        return EXIT_FAILURE;
    }
    VendorFn vendor = load_vendor();
    try {
        for (size_t k = 0; k < g.gl_pathc; k++) {
            const std::string path = g.gl_pathv[k];
            System sys;
            if (!read_system(path, n, sys)) {
                std::fprintf(stderr, "FAILED at reading %s\n", path.c_str());
                globfree(&g);
                return EXIT_FAILURE;
            }
            const std::string fname = base_name(path);
            for (int s = 16; s <= n; s *= 2) {
                tpr::hip::PTPR p(n, s);
                std::printf("%s,%s,%d,%d,%le\n", fname.c_str(), "PTPR", n, s, eval(p, sys, xt, fast));
                tpr::hip::TPR t(n, s);
                std::printf("%s,%s,%d,%d,%le\n", fname.c_str(), "TPR", n, s, eval(t, sys, xt, fast));
            }
            tpr::hip::PCR pc(n);
            std::printf("%s,%s,%d,%d,%le\n", fname.c_str(), "PCR", n, 0, eval(pc, sys, xt, fast));
            if (vendor)
                std::printf("%s,%s,%d,%d,%le\n", fname.c_str(), "REFERENCE_ROCSPARSE", n, 0,
                            eval_vendor(vendor, sys, xt));
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        globfree(&g);
        return EXIT_FAILURE;
    }
    globfree(&g);
    return 0;
}
