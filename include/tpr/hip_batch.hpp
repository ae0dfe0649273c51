// hip_batch.hpp -- C++ RAII wrappers (header-only, over the C ABI of
// tpr_hip.h) for the engine's batch and serving surfaces, which have no
// counterpart in the reference:
//
//   tpr::hip::UnevenPlan plan(sizes, TPRH_PTPR, TPRH_ARITH_FAST);   // tprh_uneven_plan_*
//   plan.solve(d_a, d_b, d_c, d_d, d_x, stream);                      // one persistent launch
//
//   tpr::hip::Server srv({TPRH_PTPR, TPRH_ARITH_FAST, 64ull << 30, 4, 16384, nullptr});
//   long long m = srv.matrix_upload(n, a, b, c);
//   long long r = srv.submit(m, nrhs, d, x);
//   srv.wait(r); srv.release(r); srv.matrix_release(m);
//
// Errors throw tpr::hip::Error, as in hip_solvers.hpp.
#pragma once
#include <vector>

#include "hip_solvers.hpp"

namespace tpr {
namespace hip {

class UnevenPlan {
   public:
    // sizes[k]: equations of system k (powers of two >= 16); schedule: NULL or
    // 32 (s, E, L) triples indexed by log2(n)
    UnevenPlan(const std::vector<int>& sizes, tprh_method method, tprh_arith arith, const int* schedule = nullptr) {
        check(tprh_uneven_plan_create(&p_, method, arith, sizes.data(), (int)sizes.size(), schedule),
              "tprh_uneven_plan_create");
    }
    ~UnevenPlan() {
        if (p_) tprh_uneven_plan_destroy(p_);
    }
    UnevenPlan(const UnevenPlan&) = delete;
    UnevenPlan& operator=(const UnevenPlan&) = delete;

    // device pointers; b may be NULL (unit diagonal); asynchronous on stream
    void solve(const float* a, const float* b, const float* c, const float* d, float* x, void* stream = nullptr) const {
        check(tprh_uneven_solve(p_, a, b, c, d, x, stream), "tprh_uneven_solve");
    }
    long long offset(int k) const { return tprh_uneven_offset(p_, k); }
    long long equations() const { return tprh_uneven_equations(p_); }

   private:
    tprh_uneven_plan* p_ = nullptr;
};

class Server {
   public:
    explicit Server(const tprh_server_config& cfg) { check(tprh_server_create(&s_, &cfg), "tprh_server_create"); }
    ~Server() {
        if (s_) tprh_server_destroy(s_);
    }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    long long matrix_upload(int n, const float* a, const float* b, const float* c) {
        long long m;
        check(tprh_server_matrix_upload(s_, n, a, b, c, &m), "tprh_server_matrix_upload");
        return m;
    }
    long long matrix_generate(int n, tprh_gen kind, uint64_t seed, uint64_t sys) {
        long long m;
        check(tprh_server_matrix_generate(s_, n, kind, seed, sys, &m), "tprh_server_matrix_generate");
        return m;
    }
    void matrix_release(long long m) { check(tprh_server_matrix_release(s_, m), "tprh_server_matrix_release"); }

    // d: nrhs * n right-hand sides (NULL: generated from rhs_seed); x: where
    // the answers go (NULL: kept in the server's arena until release)
    long long submit(long long m, int nrhs, const float* d, float* x, uint64_t rhs_seed = 0, int j0 = 0) {
        long long r;
        check(tprh_server_submit(s_, m, nrhs, d, rhs_seed, j0, x, &r), "tprh_server_submit");
        return r;
    }
    void wait(long long r, double timeout_s = 60.0) { check(tprh_server_wait(s_, r, timeout_s), "tprh_server_wait"); }
    std::vector<long long> poll(int max = 4096) {
        std::vector<long long> ids(max);
        int cnt = 0;
        check(tprh_server_poll(s_, ids.data(), max, &cnt), "tprh_server_poll");
        ids.resize(cnt);
        return ids;
    }
    // sleeps until a request completes (or timeout_s passes: empty result)
    std::vector<long long> wait_any(double timeout_s, int max = 4096) {
        std::vector<long long> ids(max);
        int cnt = 0;
        check(tprh_server_wait_any(s_, ids.data(), max, &cnt, timeout_s), "tprh_server_wait_any");
        ids.resize(cnt);
        return ids;
    }
    tprh_request_info info(long long r) {
        tprh_request_info inf;
        check(tprh_server_request_info(s_, r, &inf), "tprh_server_request_info");
        return inf;
    }
    void copy_result(long long r, float* dst) { check(tprh_server_copy_result(s_, r, dst), "tprh_server_copy_result"); }
    void release(long long r) { check(tprh_server_request_release(s_, r), "tprh_server_request_release"); }
    tprh_server_stats stats() {
        tprh_server_stats st;
        check(tprh_server_stats_get(s_, &st), "tprh_server_stats_get");
        return st;
    }
    double now() const { return tprh_server_now(s_); }

   private:
    tprh_server* s_ = nullptr;
};

}  // namespace hip
}  // namespace tpr
