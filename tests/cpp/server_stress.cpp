// server_stress.cpp -- the serving loop's host code (tpr_amd/csrc/server.hip)
// on the CPU, against the fake HIP runtime of fakehip/, under many client
// threads.  Built with -fsanitize=thread by tests/test_server_host_stress.py.
//
// The fake device: every stream is a thread running its operations in order
// (with random sleeps, so batches on different slots overlap and finish out
// of order); events complete when their stream reaches them.  The kernel
// launchers the serving loop calls are replaced by stubs with a checkable
// rule: the "solve" writes x = d + b (b the matrix's diagonal), the
// "generator" writes small integers from (seed, system, index), so every
// answer is an exact float the client recomputes.  What is tested: the table
// plumbing (offsets, counts, jobs, the zeroed work counter), admission under
// a tight budget, generated-matrix hand-over between slots, retained and
// copied-out answers, poll / wait_any / release from concurrent threads, and
// (under ThreadSanitizer) that every host access is ordered by the server's
// locks and the events it waits on.
//
// With a fourth argument k > 0, every k-th stub launch (serve, split or
// generation) returns an error instead of queueing its work: the requests of
// that batch must fail with TPRH_ERROR_HIP, their memory must come back, and
// every other request must still be answered correctly.
//
//   server_stress <client threads> <requests per client> <seed> [fail every k-th launch]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tpr_hip.h"
#include "../../tpr_amd/csrc/launch.hpp"
#include "../../tpr_amd/csrc/launch_gen.hpp"
#include "../../tpr_amd/csrc/launch_serve.hpp"
#include "../../tpr_amd/csrc/launch_split.hpp"

static std::atomic<int> g_failures{0};
static int g_fail_every = 0;                  // 0: no injected launch failures
static std::atomic<long long> g_launches{0};  // stub launches so far

static bool inject_failure() { return g_fail_every > 0 && (g_launches.fetch_add(1) + 1) % g_fail_every == 0; }
static std::mutex g_print;

#define FAIL(...)                                      \
    do {                                               \
        std::lock_guard<std::mutex> g_(g_print);       \
        std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
        std::fprintf(stderr, __VA_ARGS__);             \
        std::fprintf(stderr, "\n");                    \
        g_failures++;                                  \
    } while (0)

// ---------------------------------------------------------------- fake HIP

struct FakeEvent {
    std::mutex mu;
    std::condition_variable cv;
    unsigned long long recorded = 0;   // generation of the latest record
    unsigned long long completed = 0;  // latest generation its stream has reached
};

struct FakeStream {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    bool running = false, stop = false;
    std::thread th;
    void loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [this] { return stop || !q.empty(); });
            if (q.empty()) return;
            std::function<void()> op = std::move(q.front());
            q.pop_front();
            running = true;
            lk.unlock();
            op();
            lk.lock();
            running = false;
            cv.notify_all();
        }
    }
    void push(std::function<void()> op) {
        std::lock_guard<std::mutex> g(mu);
        q.push_back(std::move(op));
        cv.notify_all();
    }
};

static void random_pause(int max_us) {
    thread_local std::mt19937 rng(std::hash<std::thread::id>()(std::this_thread::get_id()));
    std::this_thread::sleep_for(std::chrono::microseconds(rng() % (unsigned)(max_us + 1)));
}

hipError_t hipGetDevice(int* dev) {
    *dev = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* value, hipDeviceAttribute_t, int) {
    *value = 4;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "fake HIP error"; }
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }

hipError_t hipMalloc(void** p, size_t bytes) {
    *p = std::aligned_alloc(256, (bytes + 255) & ~(size_t)255);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void* p) { return hipFree(p); }

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    FakeStream* st = new FakeStream();
    st->th = std::thread([st] { st->loop(); });
    *s = st;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [s] { return s->q.empty() && !s->running; });
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    s->th.join();  // drains the queue first
    delete s;
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
    *e = new FakeEvent();
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    unsigned long long gen;
    {
        std::lock_guard<std::mutex> g(e->mu);
        gen = ++e->recorded;
    }
    s->push([e, gen] {
        std::lock_guard<std::mutex> g(e->mu);
        e->completed = std::max(e->completed, gen);
        e->cv.notify_all();
    });
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
    std::lock_guard<std::mutex> g(e->mu);
    return e->completed >= e->recorded ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    std::unique_lock<std::mutex> lk(e->mu);
    const unsigned long long target = e->recorded;
    e->cv.wait(lk, [e, target] { return e->completed >= target; });
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    unsigned long long target;
    {
        std::lock_guard<std::mutex> g(e->mu);
        target = e->recorded;
    }
    s->push([e, target] {
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv.wait(lk, [e, target] { return e->completed >= target; });
    });
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
    s->push([dst, src, bytes] {
        random_pause(20);
        std::memcpy(dst, src, bytes);
    });
    return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
    std::memcpy(dst, src, bytes);
    return hipSuccess;
}

// ------------------------------------------------------ the stub "device"

// Values the stub generator writes: small integers, exact in fp32.
static float coef_b(int kind, uint64_t seed, uint64_t sys, long long i) {
    return (float)((kind * 5 + seed * 3 + sys * 11 + (uint64_t)i) % 997 + 1);
}
static float rhs_val(uint64_t seed, uint64_t sys, long long j, long long i) {
    return (float)((seed * 7 + sys * 13 + (uint64_t)j * 31 + (uint64_t)i) % 1009);
}

static thread_local std::string t_err;

namespace tprh {

tprh_status set_error(tprh_status st, const std::string& msg) {
    t_err = msg;
    return st;
}
tprh_status set_hip_error(hipError_t, const char* what) {
    t_err = what;
    return TPRH_ERROR_HIP;
}

// The library's default slice shapes: s = min(n, 256), one DPP row per slice.
tprh_status class_shape(int n, const int*, int* s, int* E, int* L, bool allow_split, bool* split, int, int) {
    if (n < 16 || (n & (n - 1))) return set_error(TPRH_ERROR_INVALID_VALUE, "n must be a power of two >= 16");
    *s = std::min(n, 256);
    if (!default_shape(*s, E, L)) return set_error(TPRH_ERROR_INVALID_VALUE, "bad s");
    *split = (long long)(n / *s) * *L > 1024;
    if (*split && !allow_split) return set_error(TPRH_ERROR_NOT_SUPPORTED, "split");
    return TPRH_SUCCESS;
}

hipError_t serve_occupancy(int, int, size_t, int* blocks) {
    *blocks = 1;
    return hipSuccess;
}

static int shape_n(const UnevenTask& t) {
    static const int EL[][2] = {{1, 16}, {2, 16}, {4, 16}, {8, 16}, {16, 16}, {8, 32}, {4, 64}, {8, 64}};
    if (t.shape < 0 || t.shape > 7) return -1;
    const int E = EL[t.shape][0], L = EL[t.shape][1];
    if ((1 << t.lg_lanes) != t.m * L) return -1;
    return t.m * E * L;
}

hipError_t launch_serve(int, int, const ServeLaunch& p, hipStream_t stream) {
    if (inject_failure()) return hipErrorUnknown;
    const ServeLaunch q = p;
    stream->push([q] {
        random_pause(300);
        if (*q.counter != 0) FAIL("serve launch: work counter %d at start", *q.counter);
        for (int t = 0; t < q.ntasks; t++) {
            const UnevenTask& tk = q.tasks[t];
            const int n = shape_n(tk);
            if (n <= 0 || tk.count < 1 || tk.count * (1 << tk.lg_lanes) > 1024) {
                FAIL("serve launch: bad work item %d", t);
                continue;
            }
            for (int k = 0; k < tk.count; k++) {
                const ServeSys& ss = q.sys[tk.first + k];
                const float* b = q.base + ss.coef + n;
                const float* d = q.base + ss.rhs;
                float* x = q.base + ss.x;
                for (int i = 0; i < n; i++) x[i] = d[i] + b[i];
            }
        }
        *q.counter = q.ntasks;  // as the kernel leaves it: the next batch's table copy must clear it
    });
    return hipSuccess;
}

hipError_t launch_serve_split(int, int, const ServeSplitLaunch& p, int E, int L, hipStream_t stream) {
    if (inject_failure()) return hipErrorUnknown;
    const ServeSplitLaunch q = p;
    if (!serve_split_shape(E, L) || p.n % (E * L)) FAIL("split launch: shape %dx%d for n=%d", E, L, p.n);
    const int m = p.n / (E * L);
    stream->push([q, m] {
        random_pause(300);
        for (int k = 0; k < q.nsys; k++) {
            const ServeSplitSys& ss = q.sys[k];
            const float* b = q.base + ss.coef + q.n;
            const float* d = q.base + ss.rhs;
            float* x = q.base + ss.x;
            float* rows = q.base + ss.rows;
            float* xed = q.base + ss.xed;
            for (int j = 0; j < m * kSplitRowWords; j++) rows[j] = 0.0f;  // the workspace is written too
            for (int j = 0; j < m; j++) xed[j] = 0.0f;
            for (int i = 0; i < q.n; i++) x[i] = d[i] + b[i];
        }
    });
    return hipSuccess;
}

hipError_t launch_generate_jobs(const GenJob* jobs, int njobs, long long total, float* base, hipStream_t stream) {
    if (inject_failure()) return hipErrorUnknown;
    stream->push([jobs, njobs, total, base] {
        random_pause(100);
        long long at = 0;
        for (int k = 0; k < njobs; k++) {
            const GenJob& jb = jobs[k];
            const long long n = 1LL << jb.lg_n;
            if (jb.first != at) FAIL("generation job %d starts at %lld, expected %lld", k, jb.first, at);
            float* dst = base + jb.dst;
            if (jb.coef_kind == 0) {
                for (int j = 0; j < jb.nrhs; j++)
                    for (long long i = 0; i < n; i++) dst[j * n + i] = rhs_val(jb.seed, jb.sys, jb.j0 + j, i);
                at += n * jb.nrhs;
            } else {
                for (long long i = 0; i < n; i++) {
                    dst[i] = -1.0f;
                    dst[n + i] = coef_b(jb.coef_kind - 1, jb.seed, jb.sys, i);
                    dst[2 * n + i] = -1.0f;
                }
                at += n;
            }
        }
        if (at != total) FAIL("generation launch covers %lld values, expected %lld", at, total);
    });
    return hipSuccess;
}

}  // namespace tprh

// ------------------------------------------------------------ the clients

struct Mat {
    long long id;
    int n;
    bool generated;
    int kind;
    uint64_t seed, sys;
    std::vector<float> b;  // uploaded
    float bval(long long i) const { return generated ? coef_b(kind, seed, sys, i) : b[i]; }
};

static void check_ok(tprh_status st, const char* what) {
    if (st != TPRH_SUCCESS) FAIL("%s: status %d (%s)", what, (int)st, t_err.c_str());
}

static void client(tprh_server* sv, int tid, int iters, uint64_t seed, std::atomic<long long>* solved,
                   std::atomic<long long>* failed) {
    std::mt19937_64 rng(seed * 1000 + tid);
    const int sizes[] = {16, 64, 256, 1024, 4096, 16384, 32768, 65536};
    for (int it = 0; it < iters; it++) {
        Mat mt;
        mt.n = sizes[rng() % 8];
        mt.generated = rng() % 2;
        mt.sys = (uint64_t)tid * 100000 + it;
        tprh_status st;
        if (mt.generated) {
            mt.kind = (int)(rng() % 4);
            mt.seed = rng() % 1000;
            st = tprh_server_matrix_generate(sv, mt.n, (tprh_gen)mt.kind, mt.seed, mt.sys, &mt.id);
        } else {
            mt.sys = 0;  // an uploaded matrix's generator id: synthetic right-hand sides use 0
            std::vector<float> a(mt.n, -1.0f), c(mt.n, -1.0f);
            mt.b.resize(mt.n);
            for (int i = 0; i < mt.n; i++) mt.b[i] = (float)(rng() % 500 + 1);
            st = tprh_server_matrix_upload(sv, mt.n, a.data(), mt.b.data(), c.data(), &mt.id);
        }
        if (st == TPRH_ERROR_BUDGET) {
            std::this_thread::sleep_for(std::chrono::microseconds(200));
            continue;  // other clients hold the arena: try another matrix
        }
        check_ok(st, "matrix");
        if (st != TPRH_SUCCESS) continue;
        // one to three requests on the matrix, issued before any is waited for
        struct Req {
            long long id;
            int nrhs, j0;
            uint64_t rhs_seed;
            std::vector<float> d, x;
            bool host_d, host_x;
        };
        std::vector<Req> reqs(1 + rng() % 3);
        bool released = false;
        for (Req& r : reqs) {
            r.nrhs = 1 + (int)(rng() % (mt.n >= 16384 ? 2 : 8));
            r.j0 = (int)(rng() % 50);
            r.rhs_seed = rng() % 1000;
            r.host_d = rng() % 2;
            r.host_x = rng() % 2;
            const size_t cnt = (size_t)r.nrhs * mt.n;
            if (r.host_d) {
                r.d.resize(cnt);
                for (size_t i = 0; i < cnt; i++) r.d[i] = (float)(rng() % 700);
            }
            if (r.host_x) r.x.assign(cnt, -7.0f);
            st = tprh_server_submit(sv, mt.id, r.nrhs, r.host_d ? r.d.data() : nullptr, r.rhs_seed, r.j0,
                                    r.host_x ? r.x.data() : nullptr, &r.id);
            check_ok(st, "submit");
            if (st != TPRH_SUCCESS) r.id = -1;
            if (!released && rng() % 4 == 0) {  // the matrix outlives its release while requests use it
                check_ok(tprh_server_matrix_release(sv, mt.id), "matrix release (early)");
                released = true;
            }
            if (released) break;
        }
        for (Req& r : reqs) {
            if (r.id <= 0) continue;
            // every other request is waited for in short slices, so that
            // some waits end at their deadline while the request completes
            // or fails: such a wait must report the request's own outcome
            // once it has one, never a timeout for a finished request
            tprh_status ws;
            if (rng() % 2) {
                do {
                    ws = tprh_server_wait(sv, r.id, 1e-5 * (double)(rng() % 50));
                } while (ws == TPRH_ERROR_TIMEOUT);
            } else {
                ws = tprh_server_wait(sv, r.id, 120.0);
            }
            tprh_request_info inf;
            check_ok(tprh_server_request_info(sv, r.id, &inf), "info");
            if (ws == TPRH_ERROR_HIP && g_fail_every > 0) {  // its batch's launch was made to fail
                if (inf.state != 3 || inf.x) FAIL("failed request %lld: state %d", r.id, inf.state);
                check_ok(tprh_server_request_release(sv, r.id), "release (failed)");
                failed->fetch_add(1);
                continue;
            }
            check_ok(ws, "wait");
            if (inf.state != 2) FAIL("request %lld state %d after wait", r.id, inf.state);
            std::vector<float> got;
            if (r.host_x) {
                got = r.x;
                if (inf.x) FAIL("request %lld retained answers it copied out", r.id);
            } else {
                if (!inf.x) FAIL("request %lld holds no answers", r.id);
                got.assign((size_t)r.nrhs * mt.n, -7.0f);
                check_ok(tprh_server_copy_result(sv, r.id, got.data()), "copy_result");
            }
            for (int j = 0; j < r.nrhs; j++) {
                for (int i = 0; i < mt.n; i++) {
                    const size_t k = (size_t)j * mt.n + i;
                    const float d = r.host_d ? r.d[k] : rhs_val(r.rhs_seed, mt.sys, r.j0 + j, i);
                    const float want = d + mt.bval(i);
                    if (got[k] != want) {
                        FAIL("request %lld (n=%d nrhs=%d, %s matrix, %s d, %s x) rhs %d eq %d: %g, expected %g",
                             r.id, mt.n, r.nrhs, mt.generated ? "generated" : "uploaded",
                             r.host_d ? "host" : "generated", r.host_x ? "host" : "retained", j, i, got[k], want);
                        j = r.nrhs;
                        break;
                    }
                }
            }
            check_ok(tprh_server_request_release(sv, r.id), "release");
            solved->fetch_add(r.nrhs);
        }
        if (!released) check_ok(tprh_server_matrix_release(sv, mt.id), "matrix release");
    }
}

int main(int argc, char** argv) {
    const int threads = argc > 1 ? std::atoi(argv[1]) : 4;
    const int iters = argc > 2 ? std::atoi(argv[2]) : 40;
    const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1;
    g_fail_every = argc > 4 ? std::atoi(argv[4]) : 0;

    tprh_server_config cfg{};
    cfg.method = TPRH_PTPR;
    cfg.arith = TPRH_ARITH_FAST;
    cfg.hbm_budget = (size_t)12 << 20;  // tight: requests wait for memory
    cfg.max_inflight = 3;
    cfg.max_batch_systems = 24;
    cfg.min_batch_systems = seed % 2 ? 8 : 0;
    cfg.max_wait_us = seed % 2 ? 300.0 : 0.0;
    cfg.adaptive_window = (seed / 2) % 2;  // seeds 2, 3: the adaptive window
    tprh_server* sv = nullptr;
    check_ok(tprh_server_create(&sv, &cfg), "create");
    if (!sv) return 2;

    std::atomic<long long> solved{0}, failed{0};
    std::atomic<bool> done{false};
    // a thread that drains the completion list the whole time (ids it sees
    // may already be released: it only counts them) and one that reads stats
    std::thread poller([&] {
        std::vector<long long> ids(64);
        while (!done.load()) {
            int cnt = 0;
            check_ok(tprh_server_wait_any(sv, ids.data(), (int)ids.size(), &cnt, 0.002), "wait_any");
        }
    });
    std::thread stats([&] {
        while (!done.load()) {
            tprh_server_stats s;
            check_ok(tprh_server_stats_get(sv, &s), "stats");
            if (s.bytes_in_use > s.budget) FAIL("arena use %zu above the budget %zu", s.bytes_in_use, s.budget);
            std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
    });
    std::vector<std::thread> cl;
    for (int t = 0; t < threads; t++) cl.emplace_back(client, sv, t, iters, seed, &solved, &failed);
    for (auto& t : cl) t.join();
    done = true;
    poller.join();
    stats.join();

    tprh_server_stats s;
    check_ok(tprh_server_stats_get(sv, &s), "stats");
    if (s.failed != failed.load()) FAIL("%lld requests failed, clients saw %lld", s.failed, failed.load());
    if (g_fail_every == 0 && s.failed != 0) FAIL("%lld requests failed", s.failed);
    if (g_fail_every > 0 && s.failed == 0) FAIL("no injected failure reached a request");
    if (s.completed + s.failed != s.submitted)
        FAIL("%lld submitted, %lld completed, %lld failed", s.submitted, s.completed, s.failed);
    if (s.bytes_in_use != 0) FAIL("%zu arena bytes still in use after every release", s.bytes_in_use);
    if (s.systems != solved.load()) FAIL("server solved %lld systems, clients checked %lld", s.systems, solved.load());
    check_ok(tprh_server_destroy(sv), "destroy");
    std::printf("server_stress: %d clients, %lld requests (%lld failed), %lld systems, %lld batches, max in flight %d, "
                "%s\n",
                threads, s.completed, s.failed, s.systems, s.batches, s.max_inflight_seen,
                g_failures.load() ? "FAILED" : "ok");
    return g_failures.load() ? 1 : 0;
}
