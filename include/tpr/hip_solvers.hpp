// hip_solvers.hpp -- C++ mirror of the reference's solver classes on the
// MI355X engine (header-only, over the C ABI of tpr_hip.h).
//
//   tpr::hip::PTPR t(n, s);                       // src/ptpr.hpp:58
//   t.set_tridiagonal_system(a, c, rhs);          // 3-arg, unit diagonal (src/ptpr.cpp:39-43)
//   t.set_tridiagonal_system(a, diag, c, rhs);    // 4-arg, divides by diag (src/ptpr.hpp:80-90)
//   int flops = t.solve();                        // src/ptpr.cpp:103-125
//   t.get_ans(x);                                 // src/ptpr.cpp:362-367
//
// The same surface for TPR (src/tpr.hpp:53-97), PCR (src/pcr.hpp) and CR
// (src/cr.hpp), and the CUDA free functions as tpr::hip::ptpr_cu & co.
// (src/ptpr.cuh:50-57, src/tpr.cuh:50-57).  Differences from the reference:
//   * errors throw tpr::hip::Error (the reference prints and exits);
//   * the caller's arrays are read, never modified (the reference divides
//     them by diag in place and overwrites them during solve);
//   * the arithmetic defaults to the exact mode, whose results are
//     bit-identical to the reference's CPU classes; set_arith(FAST) selects
//     FMA and the hardware reciprocal.
#pragma once
#include <stdexcept>
#include <string>

#include "../tpr_hip.h"

namespace tpr {
namespace hip {

class Error : public std::runtime_error {
   public:
    Error(tprh_status st, const char* what)
        : std::runtime_error(std::string(what) + ": " + tprh_last_error() + " (" + tprh_status_string(st) + ")"),
          status(st) {}
    tprh_status status;
};

inline void check(tprh_status st, const char* what) {
    if (st != TPRH_SUCCESS) throw Error(st, what);
}

class SolverHandle {
   public:
    SolverHandle(tprh_method method, int n, int s) : n_(n), s_(s), method_(method) {
        check(tprh_solver_create(&h_, method, n, s), "tprh_solver_create");
    }
    ~SolverHandle() {
        if (h_) tprh_solver_destroy(h_);
    }
    // a copy is a fresh solver of the same size (src/ptpr.hpp:74-78)
    SolverHandle(const SolverHandle& o) : SolverHandle(o.method_, o.n_, o.s_) {}
    SolverHandle& operator=(const SolverHandle&) = delete;

    void set_arith(tprh_arith arith) { check(tprh_solver_set_arith(h_, arith), "tprh_solver_set_arith"); }
    void set_tridiagonal_system(float* a, float* c, float* rhs) {
        check(tprh_solver_set_system(h_, a, nullptr, c, rhs), "set_tridiagonal_system");
    }
    void set_tridiagonal_system(float* a, float* diag, float* c, float* rhs) {
        check(tprh_solver_set_system(h_, a, diag, c, rhs), "set_tridiagonal_system");
    }
    int solve() {
        int flops = 0;
        check(tprh_solver_solve(h_, &flops), "solve");
        return flops;
    }
    int get_ans(float* x) {
        check(tprh_solver_get_ans(h_, x), "get_ans");
        return 0;
    }

   protected:
    int n_, s_;
    tprh_method method_;
    tprh_solver* h_ = nullptr;
};

class PTPR : public SolverHandle {
   public:
    PTPR(int n, int s) : SolverHandle(TPRH_PTPR, n, s) {}
    // src/ptpr.hpp:53-56: the 6-argument constructor uses the dividing setter
    PTPR(float* a, float* diag, float* c, float* rhs, int n, int s) : SolverHandle(TPRH_PTPR, n, s) {
        set_tridiagonal_system(a, diag, c, rhs);
    }
};

class TPR : public SolverHandle {
   public:
    TPR(int n, int s) : SolverHandle(TPRH_TPR, n, s) {}
    // src/tpr.hpp:53-56: unlike PTPR, TPR's 6-argument constructor calls the
    // 3-argument setter, so diag is ignored (a unit diagonal is assumed)
    TPR(float* a, float* diag, float* c, float* rhs, int n, int s) : SolverHandle(TPRH_TPR, n, s) {
        (void)diag;
        set_tridiagonal_system(a, c, rhs);
    }
};

class PCR : public SolverHandle {
   public:
    explicit PCR(int n) : SolverHandle(TPRH_PCR, n, 0) {}
    // src/pcr.hpp:14-17 (the 5-argument constructor stores a, c, rhs; the
    // division by diag is done by the 4-argument setter only)
    PCR(float* a, float* diag, float* c, float* rhs, int n) : SolverHandle(TPRH_PCR, n, 0) {
        (void)diag;
        set_tridiagonal_system(a, c, rhs);
    }
};

class CR : public SolverHandle {
   public:
    explicit CR(int n) : SolverHandle(TPRH_CR, n, 0) {}
    // CR's 4-argument setter stores a, c, rhs and ignores diag
    // (src/cr.cpp:17-21, reached from the constructor src/cr.hpp:31-34); so
    // does the C ABI for TPRH_CR
    CR(float* a, float* diag, float* c, float* rhs, int n) : SolverHandle(TPRH_CR, n, 0) {
        (void)diag;
        set_tridiagonal_system(a, c, rhs);
    }
};

// The CUDA free functions (unit diagonal, host pointers, synchronous).
inline void ptpr_cu(float* a, float* c, float* rhs, float* x, int n, int s) {
    check(tprh_ptpr_cu(a, c, rhs, x, n, s), "ptpr_cu");
}
inline void tpr_cu(float* a, float* c, float* rhs, float* x, int n, int s) {
    check(tprh_tpr_cu(a, c, rhs, x, n, s), "tpr_cu");
}
inline void pcr_cu(float* a, float* c, float* rhs, float* x, int n) { check(tprh_pcr_cu(a, c, rhs, x, n), "pcr_cu"); }
inline void cr_cu(float* a, float* c, float* rhs, float* x, int n) { check(tprh_cr_cu(a, c, rhs, x, n), "cr_cu"); }

}  // namespace hip
}  // namespace tpr
