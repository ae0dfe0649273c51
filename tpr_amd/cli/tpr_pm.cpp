// tpr_pm -- benchmark driver of the reference (src/pm.cpp:118-268) for the
// MI355X engine:
//
//     tpr_pm n s iter solver [destination for the result]
//
// Solvers (case-insensitive, as src/pm.cpp:66-94):
//   cuptpr, hipptpr   PTPR on the GPU   (PTPR_CU::ptpr_cu)
//   cutpr,  hiptpr    TPR on the GPU    (TPR_CU::tpr_cu)
//   hippcr            PCR on the GPU    (PTPR_CU::pcr_cu), n <= 4096
//   hipcr             CR on the GPU     (TPR_CU::cr_cu),   n <= 4096
//   cusparse, rocsparse  the vendor slot (REFERENCE_CUSPARSE, src/pm.cpp:240-246):
//                     rocSPARSE's gtsv without pivoting, from libtpr_rocsparse.so
// The reference spellings cuptpr / cutpr / cusparse are names for the HIP and
// rocSPARSE paths; there is no CUDA code behind them.  The reference's CPU
// solvers (tpr, pcr, ptpr, lapack) are not part of this GPU engine.
//
// Protocol of src/pm.cpp:226-239: per iteration the input is regenerated
// (ExampleRandomRHSInput, src/system.cpp:146-158: a = c = -1/6, diag = 1,
// rhs ~ U(-1,1)), outside the timed region, then solved from host arrays.
// The time is the kernel only, from device events around the launch, as the
// reference's DeviceTimer records it (src/ptpr.cu:574-587); one line
// "n,s,solver,ms" per iteration (src/pm.cpp:262-266).  The optional 5th
// argument writes x as "%f" values separated by commas (src/lib.hpp:98-127).
//
// Environment: TPR_ARITH=exact (default; bit-identical to the reference's CPU
// solver) or fast; TPR_PM_INPUT=random (default) or fixed, the reference's
// ExampleFixedInput (rhs = i + 1, src/system.cpp:118-130) whose answer for
// n = 1024 is test/ans1024.txt.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "tpr/hip_solvers.hpp"

namespace {

enum class Solver { PTPR, TPR, PCR, CR, VENDOR };

bool parse_solver(std::string name, Solver* out) {
    std::transform(name.begin(), name.end(), name.begin(), [](unsigned char ch) { return std::tolower(ch); });
    if (name == "cuptpr" || name == "hipptpr") *out = Solver::PTPR;
    else if (name == "cutpr" || name == "hiptpr") *out = Solver::TPR;
    else if (name == "hippcr") *out = Solver::PCR;
    else if (name == "hipcr") *out = Solver::CR;
    else if (name == "cusparse" || name == "rocsparse") *out = Solver::VENDOR;
    else if (name == "tpr" || name == "pcr" || name == "ptpr" || name == "lapack") {
        std::cerr << "Solver `" << name << "` is a CPU solver of the reference and is not part of this GPU engine; "
                  << "use cuptpr/hipptpr, cutpr/hiptpr, hippcr or hipcr.\n";
        std::exit(EXIT_FAILURE);
    } else {
        return false;
    }
    return true;
}

struct Input {  // ExampleRandomRHSInput (src/system.cpp:146-158) / ExampleFixedInput
    Input(int n_, bool fixed_) : n(n_), fixed(fixed_), a(n_), diag(n_), c(n_), rhs(n_), rng(0), dist(-1.0, 1.0) {}
    void assign() {
        for (int i = 0; i < n; i++) {
            a[i] = static_cast<float>(-1.0 / 6.0);
            c[i] = static_cast<float>(-1.0 / 6.0);
            diag[i] = 1.0f;
            rhs[i] = fixed ? static_cast<float>(1.0 * (i + 1)) : static_cast<float>(dist(rng));
        }
        a[0] = 0.0f;
        c[n - 1] = 0.0f;
    }
    int n;
    bool fixed;
    std::vector<float> a, diag, c, rhs;
    std::mt19937 rng;
    std::uniform_real_distribution<double> dist;
};

// The vendor slot as the reference runs it (src/reference_cusparse.cu:164-198):
// per call the inputs are copied to the device, the gtsv call alone is timed
// with device events, and x is copied back.  The library is looked up beside
// libtpr_hip.so at run time, as tpr_nstab does.
class VendorSlot {
public:
    explicit VendorSlot(int n) : n_(n) {
        Dl_info info;
        std::string dir = ".";
        if (dladdr(reinterpret_cast<void*>(&tprh_version), &info) && info.dli_fname) {
            std::string f = info.dli_fname;
            const size_t k = f.find_last_of('/');
            if (k != std::string::npos) dir = f.substr(0, k);
        }
        void* h = dlopen((dir + "/libtpr_rocsparse.so").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) throw std::runtime_error(std::string("the rocSPARSE slot needs libtpr_rocsparse.so: ") + dlerror());
        fn_ = reinterpret_cast<Fn>(dlsym(h, "tprh_vendor_gtsv_batch"));
        err_ = reinterpret_cast<const char* (*)()>(dlsym(h, "tprh_vendor_last_error"));
        if (!fn_ || !err_) throw std::runtime_error("libtpr_rocsparse.so lacks tprh_vendor_gtsv_batch");
        if (n <= 0) throw std::runtime_error("n must be positive");
        hip(hipMalloc(reinterpret_cast<void**>(&dev_), 5 * sizeof(float) * (size_t)n), "hipMalloc");
        hip(hipEventCreate(&ev_[0]), "hipEventCreate");
        hip(hipEventCreate(&ev_[1]), "hipEventCreate");
    }
    ~VendorSlot() {
        for (hipEvent_t e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (dev_) (void)hipFree(dev_);
    }
    // a, diag, c, rhs: host arrays of n; x: host; returns the gtsv time in ms
    float solve(const float* a, const float* diag, const float* c, const float* rhs, float* x) {
        const size_t nb = sizeof(float) * (size_t)n_;
        const float* src[4] = {a, diag, c, rhs};
        for (int k = 0; k < 4; k++) hip(hipMemcpy(dev_ + (size_t)k * n_, src[k], nb, hipMemcpyHostToDevice), "hipMemcpy");
        hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip(hipEventRecord(ev_[0], nullptr), "hipEventRecord");
        const int st = fn_(dev_, dev_ + n_, dev_ + 2 * (size_t)n_, dev_ + 3 * (size_t)n_, dev_ + 4 * (size_t)n_, n_, 1,
                           nullptr);
        if (st != 0) throw std::runtime_error(std::string("rocSPARSE gtsv: ") + err_());
        hip(hipEventRecord(ev_[1], nullptr), "hipEventRecord");
        hip(hipEventSynchronize(ev_[1]), "hipEventSynchronize");
        float ms = -1.0f;
        hip(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
        hip(hipMemcpy(x, dev_ + 4 * (size_t)n_, nb, hipMemcpyDeviceToHost), "hipMemcpy");
        return ms;
    }

private:
    using Fn = int (*)(const float*, const float*, const float*, const float*, float*, int, int, void*);
    static void hip(hipError_t e, const char* what) {
        if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }
    int n_;
    Fn fn_ = nullptr;
    const char* (*err_)() = nullptr;
    float* dev_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 5) {  // src/pm.cpp:120-125
        std::cerr << "Invalid command line args.\n";
        std::cerr << "Usage: " << argv[0] << " n s iter solver [destination for the result]\n";
        return EXIT_FAILURE;
    }
    const int n = std::atoi(argv[1]);
    const int s = std::atoi(argv[2]);
    const int iter = std::atoi(argv[3]);
    Solver solver;
    if (!parse_solver(argv[4], &solver)) {
        std::cerr << "Solver Not Found.\n";  // src/pm.cpp:90-93
        return EXIT_FAILURE;
    }
    const char* ar = std::getenv("TPR_ARITH");
    const tprh_arith arith = (ar && std::string(ar) == "fast") ? TPRH_ARITH_FAST : TPRH_ARITH_EXACT;

    const char* inp = std::getenv("TPR_PM_INPUT");
    Input input(n, inp && std::string(inp) == "fixed");
    std::vector<float> times;
    times.reserve(iter > 0 ? iter : 0);
    try {
        if (solver == Solver::VENDOR) {  // src/pm.cpp:240-246: the diagonal is 1 (the 3-array input)
            VendorSlot vs(n);
            std::vector<float> ones(n > 0 ? n : 0, 1.0f);
            for (int i = 0; i < iter; i++) {
                input.assign();
                times.push_back(vs.solve(input.a.data(), ones.data(), input.c.data(), input.rhs.data(),
                                         input.diag.data()));
            }
        } else {
            tpr::hip::SolverHandle* h = nullptr;
            switch (solver) {
                case Solver::PTPR: h = new tpr::hip::PTPR(n, s); break;
                case Solver::TPR: h = new tpr::hip::TPR(n, s); break;
                case Solver::PCR: h = new tpr::hip::PCR(n); break;
                default: h = new tpr::hip::CR(n); break;
            }
            h->set_arith(arith);
            tpr::hip::check(tprh_set_kernel_timing(1), "tprh_set_kernel_timing");
            for (int i = 0; i < iter; i++) {
                input.assign();
                h->set_tridiagonal_system(input.a.data(), input.c.data(), input.rhs.data());
                h->solve();
                h->get_ans(input.diag.data());  // the reference returns x in sys.diag
                times.push_back(tprh_last_kernel_ms());
            }
            delete h;
        }
    } catch (const std::exception& e) {  // tpr::hip::Error and the vendor slot's errors
        std::cerr << "[Error] " << e.what() << "\n";
        return EXIT_FAILURE;
    }

    if (argc >= 6) {  // src/pm.cpp:253-256, src/lib.hpp:98-127
        FILE* fp = std::fopen(argv[5], "w");
        if (!fp) {
            std::fprintf(stderr, "[%s][Error] Failed to open `%s`\n", __FILE__, argv[5]);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < n; i++) {
            std::fprintf(fp, "%f", input.diag[i]);
            if (i < n - 1) std::fprintf(fp, ",");
        }
        std::fprintf(fp, "\n");
        std::fclose(fp);
    }
    for (int i = 0; i < iter; i++) std::cout << n << "," << s << "," << argv[4] << "," << times[i] << "\n";
    return 0;
}
