// server.hip -- the serving loop behind tprh_server_* (include/tpr_hip.h).
//
// Host code only.  One device arena (a single hipMalloc of the budget) holds
// resident matrices and admitted requests; a first-fit allocator with
// coalescing hands out 256-byte aligned blocks, so the loop itself never
// calls hipMalloc.  A worker thread runs
//
//   1. completion: every in-flight batch whose event has fired retires its
//      requests (answers copied out or retained, arena blocks freed);
//   2. admission + launch: while a batch slot is free, queued requests are
//      taken in FIFO order as long as the arena has room for them and the
//      batch has room for their right-hand sides; the batch's tables are
//      built on the host and copied through a pinned staging buffer in one
//      transfer, and on the slot's own stream one launch generates the
//      coefficients of generated matrices that this batch uses first and the
//      synthetic right-hand sides, one persistent serve kernel solves every
//      class within a workgroup, and one split launch set per larger class
//      solves the rest;
//   3. idle: sleeps on a condition variable when nothing is queued or in
//      flight, otherwise polls the in-flight events.
//
// Batches on different slots overlap on the device.  Every stream has the
// default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/tpr_hip.h"
#include "launch.hpp"
#include "launch_gen.hpp"
#include "launch_serve.hpp"
#include "launch_split.hpp"

namespace tprh {
tprh_status set_error(tprh_status st, const std::string& msg);
tprh_status set_hip_error(hipError_t e, const char* what);
tprh_status class_shape(int n, const int* schedule, int* s, int* E, int* L, bool allow_split, bool* split,
                        int method, int arith);
}  // namespace tprh

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, cap) bytes; free blocks coalesce.
class Arena {
public:
    explicit Arena(size_t cap) : cap_(cap) {
        if (cap > 0) free_[0] = cap;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(bytes == 0 ? 1 : bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(bytes == 0 ? 1 : bytes);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        size_t start = off, len = bytes;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && next->first == off + bytes) {
            len += next->second;
            free_.erase(next);
        }
        free_[start] = len;
    }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t cap() const { return cap_; }

private:
    size_t cap_, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
};

// Generation state of a generated matrix: its coefficients are written by the
// first batch that uses it, in that batch's generation launch.
enum { kGenPending = 0, kGenRunning = 1, kGenDone = 2 };

struct Matrix {
    int n = 0;
    size_t off = 0, bytes = 0;  // packed a | b | c in the arena
    uint64_t sys = 0;           // generator id (right-hand sides generated for it use it)
    int refs = 0;               // queued or in-flight requests
    bool released = false;
    // uploaded matrices: the upload's event (complete when the upload returns)
    bool ready_known = false;   // the ready event has been seen complete
    hipEvent_t ready = nullptr;
    // generated matrices (ready == nullptr)
    int gen_kind = 0;
    uint64_t gen_seed = 0;
    int gen_state = kGenDone;
    int gen_slot = -1;          // the slot whose batch generates it (kGenRunning)
};

enum { kQueued = 0, kOnDevice = 1, kDone = 2, kFailed = 3 };

struct Request {
    long long mat = 0;
    int n = 0, nrhs = 0;
    int s = 0, E = 0, L = 0;
    bool split = false;
    const float* d_src = nullptr;
    float* x_dst = nullptr;
    uint64_t rhs_seed = 0;
    int j0 = 0;
    size_t off = 0, bytes = 0;  // arena block: d | x | split workspace
    size_t x_off = 0, ws_off = 0;
    bool holds_block = false;
    int state = kQueued;
    tprh_status status = TPRH_SUCCESS;
    double t_submit = 0, t_launch = 0, t_done = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t gen_done = nullptr;     // after the batch's generation launch
    std::vector<long long> gen_mats;   // matrices this batch generates
    void* host_stage = nullptr;  // pinned: the batch's packed tables
    void* dev_table = nullptr;   // device: the same tables
    bool busy = false;
    std::vector<long long> reqs;
    long long systems = 0, equations = 0;
    double t_launched = 0;  // when launch_batch returned (server clock)
    double predicted = 0;   // its expected time on the device, s (adaptive window)
};

// The server's lock.  Client calls and the worker each hold it for a few
// microseconds; a thread that finds it taken spins on try_lock for a while
// before it sleeps.  With std::mutex alone a contended call slept in the
// kernel at once and paid a futex wake-up of several microseconds, several
// times per request (submit, wait, info, release from the client threads,
// admission and retirement from the worker), which bounded the closed-loop
// request rate.
class ServerMutex {
  public:
    void lock() {
        for (int i = 0; i < kSpins; i++) {
            if (m_.try_lock()) return;
            relax();
        }
        m_.lock();
    }
    bool try_lock() { return m_.try_lock(); }
    void unlock() { m_.unlock(); }

  private:
    static constexpr int kSpins = 2000;
    static void relax() {
#if defined(__x86_64__) && !defined(__HIP_DEVICE_COMPILE__)
        __builtin_ia32_pause();
#else
        std::this_thread::yield();
#endif
    }
    std::mutex m_;
};
using ServerLock = std::unique_lock<ServerMutex>;
using ServerGuard = std::lock_guard<ServerMutex>;

}  // namespace

struct tprh_server {
    tprh_server_config cfg{};
    int schedule[96] = {};
    bool have_schedule = false;
    int method = 0, mode = 0;
    int device = 0;
    char* base = nullptr;  // arena base (device)
    Arena arena{0};
    std::chrono::steady_clock::time_point t0;

    ServerMutex mu;
    std::condition_variable_any cv_work, cv_done;
    std::deque<long long> queue;
    long long queued_systems = 0;  // right-hand sides of the queued requests
    // bumped (under mu) by every call that may let a queued request in: a
    // submit, a release that frees arena memory, the destructor
    std::atomic<unsigned long long> wake{0};
    std::unordered_map<long long, Request> reqs;
    std::unordered_map<long long, Matrix> mats;
    std::vector<long long> completed;
    long long next_req = 1, next_mat = 1;
    std::vector<Slot> slots;
    hipStream_t upload = nullptr;
    int grid = 1;
    size_t lds_bytes = 0;
    size_t table_bytes = 0;  // capacity of a slot's packed tables (the work counter included)
    std::thread worker;
    bool stop = false;
    std::string worker_error;

    tprh_server_stats st{};
    // adaptive batching window (worker thread only): device time per
    // equation of a batch, and host time to build and launch one, as
    // exponentially weighted averages over the batches seen so far
    double s_per_eq = 0.0;
    double launch_s = 50e-6;
    // and (under mu, from the submits) time between request arrivals
    double arr_gap = 0.0;
    double t_last_submit = -1.0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// LDS of the serve kernel: a work item holds 1024 / (m L) systems of 13 m
// floats each, at most 13 * 1024 / 16 floats for any shape of the menu.
constexpr size_t kServeLds = sizeof(float) * 13 * 1024 / 16;

void free_matrix_locked(tprh_server* sv, long long id) {
    auto it = sv->mats.find(id);
    if (it == sv->mats.end()) return;
    Matrix& m = it->second;
    if (m.ready) {
        (void)hipEventSynchronize(m.ready);  // its upload is finished before the block is reused
        (void)hipEventDestroy(m.ready);
    }
    sv->arena.release(m.off, m.bytes);
    sv->mats.erase(it);
}

void drop_request_ref_locked(tprh_server* sv, Request& r) {
    auto it = sv->mats.find(r.mat);
    if (it == sv->mats.end()) return;
    if (--it->second.refs == 0 && it->second.released) free_matrix_locked(sv, r.mat);
}

// Retires the batch of a slot whose event has fired (or failed).
void retire_locked(tprh_server* sv, Slot& sl, hipError_t e) {
    const double t = sv->now();
    // before the requests drop their references: the matrices are alive
    for (long long mid : sl.gen_mats) {
        auto it = sv->mats.find(mid);
        if (it == sv->mats.end()) continue;
        it->second.gen_state = e == hipSuccess ? kGenDone : kGenPending;
        it->second.gen_slot = -1;
    }
    sl.gen_mats.clear();
    for (long long id : sl.reqs) {
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) continue;
        Request& r = it->second;
        r.t_done = t;
        if (e != hipSuccess) {
            r.state = kFailed;
            r.status = TPRH_ERROR_HIP;
            sv->st.failed++;
        } else {
            r.state = kDone;
            sv->st.completed++;
        }
        // the right-hand sides and workspace always go; the answers stay only
        // when the caller asked to keep them in the arena
        const bool keep_x = r.x_dst == nullptr && e == hipSuccess;
        if (r.holds_block) {
            if (keep_x) {
                const size_t xb = sizeof(float) * (size_t)r.nrhs * r.n;
                // keep [x_off, x_off + xb): free the parts before and after it
                const size_t head = r.x_off - r.off;
                const size_t tail_off = r.x_off + align_up(xb);
                const size_t tail = r.off + r.bytes - tail_off;
                if (head) sv->arena.release(r.off, head);
                if (tail) sv->arena.release(tail_off, tail);
                r.off = r.x_off;
                r.bytes = align_up(xb);
            } else {
                sv->arena.release(r.off, r.bytes);
                r.holds_block = false;
            }
        }
        drop_request_ref_locked(sv, r);
        sv->completed.push_back(id);
    }
    if (e == hipSuccess) {
        sv->st.systems += sl.systems;
        sv->st.equations += sl.equations;
    } else {
        sv->worker_error = std::string("batch failed: ") + hipGetErrorString(e);
    }
    sl.reqs.clear();
    sl.busy = false;
}

struct ClassKey {
    int n;
    bool operator<(const ClassKey& o) const { return n > o.n; }  // largest systems first
};

// One admitted request as the launch sees it: a copy taken under the server
// lock, so that the launch itself can run without it (other threads may add
// requests and matrices meanwhile; this request's block, its matrix and the
// matrix's ready event stay valid because the request is on the device and
// holds a reference to the matrix).
struct Staged {
    Request r;
    size_t mat_off;
    uint64_t mat_sys;
    hipEvent_t ready;     // uploaded matrix not yet seen complete: wait for it
    int wait_slot;        // generated by another slot's running batch: wait for its generation
    int gen_kind;         // >= 0: this batch generates the matrix (first request to use it)
    uint64_t gen_seed;
};

// Builds and launches one batch on slot `sl` (busy, so this thread owns its
// stream and tables) from the staged requests; called without the lock.
// Every table of the batch (work items, serve-kernel entries, split
// entries, generation jobs, the serve kernel's zeroed work counter) is
// packed into the slot's pinned staging buffer and copied with one
// transfer; then, on the slot's stream: caller right-hand sides are copied
// in, synthetic ones and the coefficients of matrices first used here are
// generated (one launch, then the gen_done event that batches on other
// slots using those matrices wait for), the serve kernel runs for every
// class within a workgroup, one split launch set per larger class, and the
// answers the caller asked for are copied out.  Returns the first HIP error.
hipError_t launch_batch(tprh_server* sv, Slot& sl, const std::vector<Staged>& items) {
    float* basef = reinterpret_cast<float*>(sv->base);
    const size_t fsz = sizeof(float);
    hipError_t e = hipSuccess;
    std::map<ClassKey, std::vector<tprh::ServeSys>> groups;          // classes within a workgroup
    std::map<int, std::pair<int, int>> shape_of;                      // n -> (E, L)
    std::map<int, std::vector<tprh::ServeSplitSys>> split_groups;     // classes above a workgroup
    std::vector<tprh::GenJob> jobs;
    long long gen_total = 0;
    bool generates = false;
    unsigned waited = 0;  // slots already waited for
    for (const Staged& it : items) {
        const Request& r = it.r;
        if (it.ready && (e = hipStreamWaitEvent(sl.stream, it.ready, 0)) != hipSuccess) return e;
        if (it.wait_slot >= 0 && !(waited >> it.wait_slot & 1u)) {
            if ((e = hipStreamWaitEvent(sl.stream, sv->slots[it.wait_slot].gen_done, 0)) != hipSuccess) return e;
            waited |= 1u << it.wait_slot;
        }
        if (it.gen_kind >= 0) {
            tprh::GenJob jb{};
            jb.seed = it.gen_seed;
            jb.sys = it.mat_sys;
            jb.dst = (long long)(it.mat_off / fsz);
            jb.first = gen_total;
            jb.nrhs = 1;
            jb.lg_n = __builtin_ctz((unsigned)r.n);
            jb.coef_kind = it.gen_kind + 1;
            jobs.push_back(jb);
            gen_total += r.n;
            generates = true;
        }
        const size_t nn = (size_t)r.n;
        if (r.d_src) {
            if ((e = hipMemcpyAsync(basef + r.off / fsz, r.d_src, fsz * nn * r.nrhs, hipMemcpyDefault, sl.stream)) !=
                hipSuccess)
                return e;
        } else {
            tprh::GenJob jb{};
            jb.seed = r.rhs_seed;
            jb.sys = it.mat_sys;
            jb.dst = (long long)(r.off / fsz);
            jb.first = gen_total;
            jb.j0 = r.j0;
            jb.nrhs = r.nrhs;
            jb.lg_n = __builtin_ctz((unsigned)r.n);
            jobs.push_back(jb);
            gen_total += (long long)r.nrhs * r.n;
        }
        shape_of[r.n] = {r.E, r.L};
        if (!r.split) {
            auto& g = groups[ClassKey{r.n}];
            for (int j = 0; j < r.nrhs; j++) {
                tprh::ServeSys ss;
                ss.coef = (long long)(it.mat_off / fsz);
                ss.rhs = (long long)(r.off / fsz + (size_t)j * nn);
                ss.x = (long long)(r.x_off / fsz + (size_t)j * nn);
                g.push_back(ss);
            }
        } else {
            // the request's workspace: rows of its nrhs systems, then their x_ed
            const size_t ms = (size_t)(r.n / r.s);
            const size_t rows0 = r.ws_off / fsz;
            auto& g = split_groups[r.n];
            for (int j = 0; j < r.nrhs; j++) {
                tprh::ServeSplitSys ss;
                ss.coef = (long long)(it.mat_off / fsz);
                ss.rhs = (long long)(r.off / fsz + (size_t)j * nn);
                ss.x = (long long)(r.x_off / fsz + (size_t)j * nn);
                ss.rows = (long long)(rows0 + (size_t)j * ms * tprh::kSplitRowWords);
                ss.xed = (long long)(rows0 + (size_t)r.nrhs * ms * tprh::kSplitRowWords + (size_t)j * ms);
                g.push_back(ss);
            }
        }
    }
    std::vector<tprh::UnevenTask> tasks;
    std::vector<tprh::ServeSys> entries;
    for (auto& kv : groups) {
        const int n = kv.first.n;
        const int E = shape_of[n].first, L = shape_of[n].second;
        const int m = n / (E * L);
        const int T = m * L;
        const int per = 1024 / T;
        const std::vector<tprh::ServeSys>& v = kv.second;
        for (size_t i = 0; i < v.size(); i += per) {
            tprh::UnevenTask tk;
            tk.shape = tprh::uneven_shape_id(E, L);
            tk.m = m;
            tk.lg_lanes = __builtin_ctz((unsigned)T);
            tk.first = (int)entries.size();
            tk.count = (int)std::min<size_t>(per, v.size() - i);
            for (int q = 0; q < tk.count; q++) entries.push_back(v[i + q]);
            tasks.push_back(tk);
        }
    }
    // pack the tables: each region's used part, 256-byte aligned
    char* hs = static_cast<char*>(sl.host_stage);
    char* dt = static_cast<char*>(sl.dev_table);
    size_t at = 0;
    auto put = [&](const void* src, size_t bytes) {
        const size_t here = at;
        if (bytes) std::memcpy(hs + here, src, bytes);
        at = align_up(here + bytes);
        return here;
    };
    const size_t o_tasks = put(tasks.data(), sizeof(tprh::UnevenTask) * tasks.size());
    const size_t o_entries = put(entries.data(), sizeof(tprh::ServeSys) * entries.size());
    std::vector<std::pair<size_t, size_t>> split_at;  // (offset, count) per split class, in map order
    for (auto& kv : split_groups)
        split_at.push_back({put(kv.second.data(), sizeof(tprh::ServeSplitSys) * kv.second.size()), kv.second.size()});
    const size_t o_jobs = put(jobs.data(), sizeof(tprh::GenJob) * jobs.size());
    const int zero = 0;  // the serve kernel's work counter, cleared by this copy
    const size_t o_counter = tasks.empty() ? 0 : put(&zero, sizeof(int));
    if (at > 0 && (e = hipMemcpyAsync(dt, hs, at, hipMemcpyHostToDevice, sl.stream)) != hipSuccess) return e;
    if (!jobs.empty() &&
        (e = tprh::launch_generate_jobs(reinterpret_cast<const tprh::GenJob*>(dt + o_jobs), (int)jobs.size(),
                                        gen_total, basef, sl.stream)) != hipSuccess)
        return e;
    // batches on other slots that use a matrix generated here wait for this
    if (generates && (e = hipEventRecord(sl.gen_done, sl.stream)) != hipSuccess) return e;
    if (!tasks.empty()) {
        tprh::ServeLaunch p;
        p.tasks = reinterpret_cast<const tprh::UnevenTask*>(dt + o_tasks);
        p.ntasks = (int)tasks.size();
        p.counter = reinterpret_cast<int*>(dt + o_counter);
        p.sys = reinterpret_cast<const tprh::ServeSys*>(dt + o_entries);
        p.base = basef;
        p.grid = std::max(1, std::min(sv->grid, p.ntasks));
        p.lds_bytes = sv->lds_bytes;
        if ((e = tprh::launch_serve(sv->method, sv->mode, p, sl.stream)) != hipSuccess) return e;
    }
    size_t k = 0;
    for (auto& kv : split_groups) {
        tprh::ServeSplitLaunch p;
        p.sys = reinterpret_cast<const tprh::ServeSplitSys*>(dt + split_at[k].first);
        p.nsys = (int)split_at[k].second;
        p.n = kv.first;
        p.base = basef;
        const auto sh = shape_of[kv.first];
        if ((e = tprh::launch_serve_split(sv->method, sv->mode, p, sh.first, sh.second, sl.stream)) != hipSuccess)
            return e;
        k++;
    }
    for (const Staged& it : items) {
        const Request& r = it.r;
        if (r.x_dst &&
            (e = hipMemcpyAsync(r.x_dst, basef + r.x_off / fsz, fsz * (size_t)r.n * r.nrhs, hipMemcpyDefault,
                                sl.stream)) != hipSuccess)
            return e;
    }
    return hipEventRecord(sl.done, sl.stream);
}

// While batches run, whether the next batch should wait for more requests
// (called by the worker under the lock, queue not empty).
// Fixed window: until min_batch_systems right-hand sides are queued or the
// oldest queued request has waited max_wait_us.
// Adaptive window: until the earliest in-flight batch is expected to finish
// within the time it takes to build and launch a batch -- so the device is
// kept busy and each batch takes everything that arrived while the previous
// ones ran -- or a batch is full, or (when max_wait_us > 0) the oldest request
// has waited max_wait_us.  Expected finish times come from the measured
// device time per equation of the batches seen so far; before the first one
// has completed nothing waits.
bool batch_should_wait(tprh_server* sv, int inflight) {
    const double now = sv->now();
    const double waited_us = 1e6 * (now - sv->reqs[sv->queue.front()].t_submit);
    if (!sv->cfg.adaptive_window)
        return inflight > 0 && sv->queued_systems < sv->cfg.min_batch_systems && waited_us < sv->cfg.max_wait_us;
    if (sv->queued_systems >= sv->cfg.max_batch_systems) return false;
    const double cap_us = sv->cfg.max_wait_us > 0.0 ? sv->cfg.max_wait_us : 4e6 * sv->launch_s;
    if (waited_us >= cap_us) return false;
    // Arrivals faster than one per launch: the worker, not the device, sets
    // the pace, and a batch of whatever is queued stays small (it finishes
    // before the next decision, so nothing is in flight to wait for).  Hold
    // the batch until it has about twice the requests that arrive during
    // one launch, so the worker spends at most about half its time
    // launching.  At light load (fewer than one arrival per launch) this
    // never applies, and requests go at once.
    if (sv->arr_gap > 0.0 && sv->launch_s >= sv->arr_gap &&
        (double)sv->queue.size() < 2.0 * sv->launch_s / sv->arr_gap)
        return true;
    if (inflight == 0 || sv->s_per_eq <= 0.0) return false;
    double first_free = 1e300;
    for (const Slot& sl : sv->slots)
        if (sl.busy) first_free = std::min(first_free, sl.t_launched + sl.predicted);
    return now + sv->launch_s < first_free;
}

// The worker polls its in-flight batches without the lock (a slot's state
// changes only on this thread) and takes the lock only when a batch has
// finished, when a slot is free and a submit or a release has happened since
// its last admission pass (the wake counter), or when it is about to sleep;
// so that client calls are not starved while batches run.
void worker_main(tprh_server* sv) {
    (void)hipSetDevice(sv->device);
    const int ns = (int)sv->slots.size();
    std::vector<std::pair<int, hipError_t>> fired;
    fired.reserve(ns);
    unsigned long long seen = ~0ull;  // wake counter at the last admission pass
    bool recheck = true;              // admission must run again without a wake
    for (;;) {
        // 1. completions
        int inflight = 0, nfree = 0;
        fired.clear();
        for (int k = 0; k < ns; k++) {
            Slot& sl = sv->slots[k];
            if (!sl.busy) {
                nfree++;
                continue;
            }
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) inflight++;
            else fired.push_back({k, q});
        }
        const unsigned long long wake = sv->wake.load(std::memory_order_acquire);
        if (fired.empty() && inflight > 0 && !recheck && (nfree == 0 || wake == seen)) {
            std::this_thread::yield();
            continue;
        }
        ServerLock lk(sv->mu);
        seen = wake;
        recheck = false;
        for (const auto& f : fired) {
            Slot& sl = sv->slots[f.first];
            if (f.second == hipSuccess && sl.equations > 0) {
                // the batch's time from launch to the completion seen here
                const double rate = (sv->now() - sl.t_launched) / (double)sl.equations;
                sv->s_per_eq = sv->s_per_eq > 0.0 ? 0.8 * sv->s_per_eq + 0.2 * rate : rate;
            }
            retire_locked(sv, sl, f.second);
        }
        if (!fired.empty()) sv->cv_done.notify_all();
        // 2. admission and launch
        bool launched = false, windowed = false;
        for (Slot& sl : sv->slots) {
            if (sl.busy || sv->queue.empty()) continue;
            if (batch_should_wait(sv, inflight)) {
                recheck = true;
                windowed = true;
                break;  // let this batch grow
            }
            std::vector<long long> batch;
            long long systems = 0, equations = 0;
            while (!sv->queue.empty()) {
                const long long id = sv->queue.front();
                Request& r = sv->reqs[id];
                if (!batch.empty() && systems + r.nrhs > sv->cfg.max_batch_systems) break;
                size_t off;
                if (!sv->arena.alloc(r.bytes, &off)) break;  // FIFO: wait for memory
                r.off = off;
                r.holds_block = true;
                const size_t xb = sizeof(float) * (size_t)r.nrhs * r.n;
                r.x_off = off + align_up(xb);
                r.ws_off = r.x_off + align_up(xb);
                sv->queue.pop_front();
                sv->queued_systems -= r.nrhs;
                batch.push_back(id);
                systems += r.nrhs;
                equations += (long long)r.nrhs * r.n;
            }
            if (batch.empty()) break;
            const double t = sv->now();
            for (long long id : batch) {
                Request& r = sv->reqs[id];
                r.state = kOnDevice;
                r.t_launch = t;
            }
            sl.reqs = batch;
            sl.systems = systems;
            sl.equations = equations;
            sl.busy = true;
            const int slot = (int)(&sl - sv->slots.data());
            std::vector<Staged> items;
            items.reserve(batch.size());
            for (long long id : batch) {
                const Request& r = sv->reqs[id];
                Matrix& m = sv->mats[r.mat];
                Staged it{r, m.off, m.sys, nullptr, -1, -1, 0};
                if (m.ready) {
                    if (!m.ready_known && hipEventQuery(m.ready) == hipSuccess) m.ready_known = true;
                    if (!m.ready_known) it.ready = m.ready;
                } else if (m.gen_state == kGenPending) {
                    m.gen_state = kGenRunning;
                    m.gen_slot = slot;
                    sl.gen_mats.push_back(r.mat);
                    it.gen_kind = m.gen_kind;
                    it.gen_seed = m.gen_seed;
                } else if (m.gen_state == kGenRunning && m.gen_slot != slot) {
                    it.wait_slot = m.gen_slot;
                }
                items.push_back(it);
            }
            lk.unlock();  // the HIP calls of the launch run without the lock
            const double t_begin = sv->now();
            hipError_t e = launch_batch(sv, sl, items);
            sl.t_launched = sv->now();
            sv->launch_s = 0.8 * sv->launch_s + 0.2 * (sl.t_launched - t_begin);
            sl.predicted = sv->s_per_eq * (double)equations;
            lk.lock();
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(sl.stream);
                retire_locked(sv, sl, e);
                sv->cv_done.notify_all();
                continue;
            }
            sv->st.batches++;
            sv->st.max_batch_systems_seen = std::max(sv->st.max_batch_systems_seen, systems);
            inflight++;
            launched = true;
            sv->st.max_inflight_seen = std::max(sv->st.max_inflight_seen, inflight);
        }
        if (sv->stop && inflight == 0 && !launched && !sv->queue.empty()) {
            // shutting down with requests that can never be admitted (the
            // memory they wait for is held by matrices or retained answers)
            for (long long id : sv->queue) {
                Request& r = sv->reqs[id];
                r.state = kFailed;
                r.status = TPRH_ERROR_BUDGET;
                r.t_done = sv->now();
                drop_request_ref_locked(sv, r);
                sv->completed.push_back(id);
                sv->st.failed++;
            }
            sv->queue.clear();
            sv->queued_systems = 0;
            sv->worker_error = "server destroyed before the request could be admitted";
            sv->cv_done.notify_all();
        }
        sv->st.queued = (int)sv->queue.size();
        if (sv->stop && sv->queue.empty() && inflight == 0) break;
        // 3. idle: sleep until a submit, a release or the destructor
        if (inflight == 0 && !launched) {
            // a batch held by the window is looked at again soon
            if (windowed) sv->cv_work.wait_for(lk, std::chrono::microseconds(10));
            else sv->cv_work.wait_for(lk, std::chrono::milliseconds(50));
            recheck = true;
        }
    }
}

// A wait's deadline on the steady clock; timeouts beyond a year are taken as
// a year, so that the conversion to the clock's integer ticks cannot overflow.
std::chrono::steady_clock::time_point wait_deadline(double timeout_s) {
    const double t = std::min(timeout_s, 365.0 * 86400.0);
    return std::chrono::steady_clock::now() +
           std::chrono::duration_cast<std::chrono::steady_clock::duration>(std::chrono::duration<double>(t));
}

tprh_status validate_n(tprh_server* sv, int n, int* s, int* E, int* L, bool* split) {
    tprh_status st = tprh::class_shape(n, sv->have_schedule ? sv->schedule : nullptr, s, E, L, true, split,
                                       (int)sv->cfg.method, (int)sv->cfg.arith);
    if (st != TPRH_SUCCESS) return st;
    if (*split && !tprh::serve_split_shape(*E, *L))
        return tprh::set_error(TPRH_ERROR_NOT_SUPPORTED, "size classes above one workgroup need s >= 256 when served");
    if (!*split && tprh::uneven_shape_id(*E, *L) < 0)
        return tprh::set_error(TPRH_ERROR_NOT_SUPPORTED, "slice shape has no path in the serve kernel");
    return TPRH_SUCCESS;
}

}  // namespace

extern "C" {

tprh_status tprh_server_create(tprh_server** out, const tprh_server_config* cfg) {
    if (!out || !cfg) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    *out = nullptr;
    if (cfg->method != TPRH_PTPR && cfg->method != TPRH_TPR)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "method must be TPRH_PTPR or TPRH_TPR");
    if (cfg->arith != TPRH_ARITH_FAST && cfg->arith != TPRH_ARITH_EXACT)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "bad arith");
    if (cfg->max_inflight < 1 || cfg->max_inflight > 16)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "max_inflight must be 1..16");
    if (cfg->max_batch_systems < 1 || cfg->max_batch_systems > (1 << 20))
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "max_batch_systems must be 1..2^20");
    if (cfg->hbm_budget < kAlign) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "hbm_budget too small");
    if (cfg->min_batch_systems < 0 || !(cfg->max_wait_us >= 0.0))
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "batching window must be non-negative");
    if (cfg->adaptive_window != 0 && cfg->adaptive_window != 1)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "adaptive_window must be 0 or 1");
    if (cfg->schedule) {
        // every row must name a shape the serve kernel has (or be zero)
        for (int lg = 4; lg < 32; lg++) {
            if (cfg->schedule[3 * lg] == 0) continue;
            int s, E, L;
            bool split;
            tprh_status st =
                tprh::class_shape(1 << lg, cfg->schedule, &s, &E, &L, true, &split, (int)cfg->method, (int)cfg->arith);
            if (st != TPRH_SUCCESS) return st;
            if (!split && tprh::uneven_shape_id(E, L) < 0)
                return tprh::set_error(TPRH_ERROR_NOT_SUPPORTED, "slice shape has no path in the serve kernel");
            if (split && !tprh::serve_split_shape(E, L))
                return tprh::set_error(TPRH_ERROR_NOT_SUPPORTED,
                                       "size classes above one workgroup need s >= 256 when served");
        }
    }
    tprh_server* sv = new (std::nothrow) tprh_server();
    if (!sv) return tprh::set_error(TPRH_ERROR_OUT_OF_MEMORY, "server allocation");
    sv->cfg = *cfg;
    if (cfg->schedule) {
        std::copy(cfg->schedule, cfg->schedule + 96, sv->schedule);
        sv->have_schedule = true;
    }
    sv->cfg.schedule = nullptr;
    sv->method = cfg->method == TPRH_PTPR ? 0 : 1;
    sv->mode = cfg->arith == TPRH_ARITH_EXACT ? 1 : 0;
    sv->t0 = std::chrono::steady_clock::now();
    auto bail = [&](hipError_t e, const char* what) {
        for (Slot& sl : sv->slots) {
            if (sl.stream) (void)hipStreamDestroy(sl.stream);
            if (sl.done) (void)hipEventDestroy(sl.done);
            if (sl.gen_done) (void)hipEventDestroy(sl.gen_done);
            if (sl.host_stage) (void)hipHostFree(sl.host_stage);
            if (sl.dev_table) (void)hipFree(sl.dev_table);
        }
        if (sv->upload) (void)hipStreamDestroy(sv->upload);
        if (sv->base) (void)hipFree(sv->base);
        delete sv;
        return tprh::set_hip_error(e, what);
    };
    hipError_t e;
    if ((e = hipGetDevice(&sv->device)) != hipSuccess) return bail(e, "hipGetDevice");
    if ((e = hipMalloc(reinterpret_cast<void**>(&sv->base), cfg->hbm_budget)) != hipSuccess)
        return bail(e, "hipMalloc(server arena)");
    sv->arena = Arena(cfg->hbm_budget);
    if ((e = hipStreamCreateWithFlags(&sv->upload, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    // per-slot tables: work items (at most one per right-hand side), the
    // serve kernel's system entries, the split launches' system entries
    // (a batch has at most max_batch_systems right-hand sides, hence at most
    // that many items, entries and requests; split classes add one alignment
    // pad each, at most 32 of them), the generation jobs (two per request at
    // most) and the work counter
    sv->table_bytes = align_up(sizeof(tprh::UnevenTask) * (size_t)cfg->max_batch_systems) +
                      align_up(sizeof(tprh::ServeSys) * (size_t)cfg->max_batch_systems) +
                      align_up(sizeof(tprh::ServeSplitSys) * (size_t)cfg->max_batch_systems) + 32 * kAlign +
                      align_up(sizeof(tprh::GenJob) * 2 * (size_t)cfg->max_batch_systems) + kAlign;
    sv->slots.resize(cfg->max_inflight);
    for (Slot& sl : sv->slots) {
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess)
            return bail(e, "hipStreamCreate");
        if ((e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&sl.gen_done, hipEventDisableTiming)) != hipSuccess)
            return bail(e, "hipEventCreate");
        if ((e = hipHostMalloc(&sl.host_stage, sv->table_bytes)) != hipSuccess) return bail(e, "hipHostMalloc");
        if ((e = hipMalloc(&sl.dev_table, sv->table_bytes)) != hipSuccess) return bail(e, "hipMalloc");
        // one table copy of each size class, now: the runtime sets up its
        // copy paths on first use (the first copy of a few tens of KB took
        // about 6 ms of host time on MI355X), which would otherwise stall
        // the first batches served
        std::memset(sl.host_stage, 0, sv->table_bytes);
        for (size_t bytes = kAlign;; bytes *= 8) {
            bytes = std::min(bytes, sv->table_bytes);
            if ((e = hipMemcpyAsync(sl.dev_table, sl.host_stage, bytes, hipMemcpyHostToDevice, sl.stream)) !=
                hipSuccess)
                return bail(e, "hipMemcpyAsync(tables)");
            if (bytes == sv->table_bytes) break;
        }
        if ((e = hipStreamSynchronize(sl.stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    }
    int cus = 0, occ = 0;
    sv->lds_bytes = kServeLds;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, sv->device)) != hipSuccess ||
        (e = tprh::serve_occupancy(sv->method, sv->mode, sv->lds_bytes, &occ)) != hipSuccess)
        return bail(e, "device query");
    sv->grid = std::max(1, cus * std::max(1, occ));
    sv->st.budget = cfg->hbm_budget;
    try {
        sv->worker = std::thread(worker_main, sv);
    } catch (...) {
        return bail(hipErrorUnknown, "worker thread");
    }
    *out = sv;
    return TPRH_SUCCESS;
}

tprh_status tprh_server_destroy(tprh_server* sv) {
    if (!sv) return TPRH_SUCCESS;
    {
        ServerGuard g(sv->mu);
        sv->stop = true;
        sv->wake++;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipStreamSynchronize(sv->upload);
    for (auto& kv : sv->mats)
        if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
    for (Slot& sl : sv->slots) {
        (void)hipStreamDestroy(sl.stream);
        (void)hipEventDestroy(sl.done);
        (void)hipEventDestroy(sl.gen_done);
        (void)hipHostFree(sl.host_stage);
        (void)hipFree(sl.dev_table);
    }
    (void)hipStreamDestroy(sv->upload);
    (void)hipFree(sv->base);
    delete sv;
    return TPRH_SUCCESS;
}

// Reserves the arena block and the record of a new matrix, under the lock.
// An uploaded matrix (ev != nullptr) is then filled by the caller without the
// lock: its id is not yet known to any other call, so nothing else touches
// the record.  A generated one is filled by the first batch that uses it.
static tprh_status new_matrix(tprh_server* sv, int n, uint64_t sys, hipEvent_t ev, int gen_kind, uint64_t gen_seed,
                              long long* id, size_t* off_out) {
    ServerGuard g(sv->mu);
    int s, E, L;
    bool split;
    tprh_status st = validate_n(sv, n, &s, &E, &L, &split);
    if (st != TPRH_SUCCESS) return st;
    const size_t bytes = sizeof(float) * 3 * (size_t)n;
    size_t off = 0;
    if (!sv->arena.alloc(bytes, &off))
        return tprh::set_error(TPRH_ERROR_BUDGET, "no room for the matrix in the server's memory budget");
    const long long mid = sv->next_mat++;
    Matrix& m = sv->mats[mid];
    m.n = n;
    m.off = off;
    m.bytes = bytes;
    m.sys = sys;
    m.ready = ev;
    if (!ev) {
        m.gen_kind = gen_kind;
        m.gen_seed = gen_seed;
        m.gen_state = kGenPending;
    }
    *id = mid;
    *off_out = off;
    return TPRH_SUCCESS;
}

tprh_status tprh_server_matrix_upload(tprh_server* sv, int n, const float* a, const float* b, const float* c,
                                      long long* matrix) {
    if (!sv || !matrix) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    if (!a || !b || !c) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    hipEvent_t ev;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return tprh::set_hip_error(e, "hipEventCreate");
    size_t off;
    tprh_status st = new_matrix(sv, n, 0, ev, 0, 0, matrix, &off);
    if (st != TPRH_SUCCESS) {
        (void)hipEventDestroy(ev);
        return st;
    }
    float* dst = reinterpret_cast<float*>(sv->base + off);
    const size_t nb = sizeof(float) * (size_t)n;
    if ((e = hipMemcpyAsync(dst, a, nb, hipMemcpyDefault, sv->upload)) != hipSuccess ||
        (e = hipMemcpyAsync(dst + n, b, nb, hipMemcpyDefault, sv->upload)) != hipSuccess ||
        (e = hipMemcpyAsync(dst + 2 * (size_t)n, c, nb, hipMemcpyDefault, sv->upload)) != hipSuccess ||
        (e = hipEventRecord(ev, sv->upload)) != hipSuccess || (e = hipEventSynchronize(ev)) != hipSuccess) {
        (void)hipStreamSynchronize(sv->upload);  // copies already queued finish before the block is reused
        ServerGuard g(sv->mu);
        free_matrix_locked(sv, *matrix);
        return tprh::set_hip_error(e, "matrix upload");
    }
    ServerGuard g(sv->mu);
    sv->mats[*matrix].ready_known = true;
    return TPRH_SUCCESS;
}

tprh_status tprh_server_matrix_generate(tprh_server* sv, int n, tprh_gen kind, uint64_t seed, uint64_t sys,
                                        long long* matrix) {
    if (kind < TPRH_GEN_DOMINANT || kind > TPRH_GEN_RANDOM_RHS)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "bad kind");
    if (!sv || !matrix) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    size_t off;
    return new_matrix(sv, n, sys, nullptr, (int)kind, seed, matrix, &off);
}

tprh_status tprh_server_matrix_release(tprh_server* sv, long long matrix) {
    if (!sv) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null server");
    ServerGuard g(sv->mu);
    auto it = sv->mats.find(matrix);
    if (it == sv->mats.end() || it->second.released)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "unknown matrix");
    it->second.released = true;
    if (it->second.refs == 0) {
        free_matrix_locked(sv, matrix);
        sv->wake++;
        sv->cv_work.notify_one();  // queued requests may fit now
    }
    return TPRH_SUCCESS;
}

tprh_status tprh_server_submit(tprh_server* sv, long long matrix, int nrhs, const float* d, uint64_t rhs_seed, int j0,
                               float* x, long long* request) {
    if (!sv || !request) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    ServerLock lk(sv->mu);
    auto it = sv->mats.find(matrix);
    if (it == sv->mats.end() || it->second.released)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "unknown matrix");
    if (nrhs < 1 || nrhs > sv->cfg.max_batch_systems)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "nrhs must be 1..max_batch_systems");
    if (j0 < 0) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "j0 must be >= 0");
    Matrix& m = it->second;
    Request r;
    r.mat = matrix;
    r.n = m.n;
    r.nrhs = nrhs;
    tprh_status st = validate_n(sv, m.n, &r.s, &r.E, &r.L, &r.split);
    if (st != TPRH_SUCCESS) return st;
    const size_t xb = align_up(sizeof(float) * (size_t)nrhs * m.n);
    r.bytes = 2 * xb + (r.split ? align_up(tprh::split_workspace_bytes(m.n / r.s, nrhs)) : 0);
    if (r.bytes > sv->arena.cap())
        return tprh::set_error(TPRH_ERROR_BUDGET, "the request needs more memory than the server's whole budget");
    r.d_src = d;
    r.x_dst = x;
    r.rhs_seed = rhs_seed;
    r.j0 = j0;
    r.t_submit = sv->now();
    if (sv->t_last_submit >= 0.0) {
        const double gap = r.t_submit - sv->t_last_submit;
        sv->arr_gap = sv->arr_gap > 0.0 ? 0.9 * sv->arr_gap + 0.1 * gap : gap;
    }
    sv->t_last_submit = r.t_submit;
    const long long id = sv->next_req++;
    sv->reqs[id] = r;
    sv->queue.push_back(id);
    sv->queued_systems += nrhs;
    sv->wake++;
    m.refs++;
    sv->st.submitted++;
    *request = id;
    lk.unlock();
    sv->cv_work.notify_one();
    return TPRH_SUCCESS;
}

tprh_status tprh_server_request_info(tprh_server* sv, long long request, tprh_request_info* info) {
    if (!sv || !info) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    ServerGuard g(sv->mu);
    auto it = sv->reqs.find(request);
    if (it == sv->reqs.end()) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "unknown request");
    const Request& r = it->second;
    info->state = r.state;
    info->n = r.n;
    info->nrhs = r.nrhs;
    info->t_submit = r.t_submit;
    info->t_launch = r.t_launch;
    info->t_done = r.t_done;
    info->x = (r.state == kDone && r.holds_block) ? reinterpret_cast<const float*>(sv->base + r.x_off) : nullptr;
    return TPRH_SUCCESS;
}

tprh_status tprh_server_wait(tprh_server* sv, long long request, double timeout_s) {
    if (!sv) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null server");
    if (!(timeout_s >= 0.0)) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "timeout must be >= 0");
    ServerLock lk(sv->mu);
    const auto deadline = wait_deadline(timeout_s);
    for (;;) {
        auto it = sv->reqs.find(request);
        if (it == sv->reqs.end()) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "unknown request");
        if (it->second.state == kDone) return TPRH_SUCCESS;
        if (it->second.state == kFailed) return tprh::set_error(it->second.status, sv->worker_error);
        if (sv->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
            // a request that completed or failed right at the deadline
            // reports its own outcome, not the timeout
            auto jt = sv->reqs.find(request);
            if (jt != sv->reqs.end() && jt->second.state == kDone) return TPRH_SUCCESS;
            if (jt != sv->reqs.end() && jt->second.state == kFailed)
                return tprh::set_error(jt->second.status, sv->worker_error);
            return tprh::set_error(TPRH_ERROR_TIMEOUT, "timed out waiting for the request");
        }
    }
}

// Hands out up to max completed ids, oldest first; ids of requests already
// released are dropped on the way.
static void take_completed_locked(tprh_server* sv, long long* ids, int max, int* count) {
    int k = 0;
    size_t i = 0;
    for (; i < sv->completed.size() && k < max; i++) {
        const long long id = sv->completed[i];
        if (sv->reqs.count(id)) ids[k++] = id;
    }
    sv->completed.erase(sv->completed.begin(), sv->completed.begin() + (long)i);
    *count = k;
}

tprh_status tprh_server_poll(tprh_server* sv, long long* ids, int max, int* count) {
    if (!sv || !count || (max > 0 && !ids)) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    ServerGuard g(sv->mu);
    take_completed_locked(sv, ids, max, count);
    return TPRH_SUCCESS;
}

tprh_status tprh_server_wait_any(tprh_server* sv, long long* ids, int max, int* count, double timeout_s) {
    if (!sv || !count || max < 1 || !ids) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument or max < 1");
    if (!(timeout_s >= 0.0)) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "timeout must be >= 0");
    ServerLock lk(sv->mu);
    const auto deadline = wait_deadline(timeout_s);
    *count = 0;
    // the list may hold only ids released meanwhile: then wait on
    while (sv->cv_done.wait_until(lk, deadline, [sv] { return !sv->completed.empty(); })) {
        take_completed_locked(sv, ids, max, count);
        if (*count > 0) break;
    }
    return TPRH_SUCCESS;
}

tprh_status tprh_server_copy_result(tprh_server* sv, long long request, float* dst) {
    if (!sv || !dst) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    // under the lock, so that a concurrent tprh_server_request_release cannot
    // free the block while it is copied
    ServerGuard g(sv->mu);
    auto it = sv->reqs.find(request);
    if (it == sv->reqs.end()) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "unknown request");
    const Request& r = it->second;
    if (r.state != kDone || !r.holds_block)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "the request holds no answers");
    const void* src = sv->base + r.x_off;
    const size_t bytes = sizeof(float) * (size_t)r.n * r.nrhs;
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDefault);
    if (e != hipSuccess) return tprh::set_hip_error(e, "hipMemcpy(result)");
    return TPRH_SUCCESS;
}

tprh_status tprh_server_request_release(tprh_server* sv, long long request) {
    if (!sv) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null server");
    ServerGuard g(sv->mu);
    auto it = sv->reqs.find(request);
    if (it == sv->reqs.end()) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "unknown request");
    Request& r = it->second;
    if (r.state != kDone && r.state != kFailed)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "the request has not completed");
    if (r.holds_block) {
        sv->arena.release(r.off, r.bytes);
        sv->wake++;
        sv->cv_work.notify_one();  // queued requests may fit now
    }
    sv->reqs.erase(it);
    // clients that wait and release without polling leave their ids in the
    // completion list: drop the released ones once the list outgrows the live
    // requests (amortised O(1) per release)
    if (sv->completed.size() > 2 * sv->reqs.size() + 4096)
        sv->completed.erase(std::remove_if(sv->completed.begin(), sv->completed.end(),
                                           [sv](long long id) { return sv->reqs.count(id) == 0; }),
                            sv->completed.end());
    return TPRH_SUCCESS;
}

double tprh_server_now(const tprh_server* sv) { return sv ? sv->now() : 0.0; }

tprh_status tprh_server_stats_get(tprh_server* sv, tprh_server_stats* out) {
    if (!sv || !out) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null argument");
    ServerGuard g(sv->mu);
    *out = sv->st;
    out->queued = (int)sv->queue.size();
    out->bytes_in_use = sv->arena.used();
    out->peak_bytes = sv->arena.peak();
    out->budget = sv->arena.cap();
    return TPRH_SUCCESS;
}

}  // extern "C"
