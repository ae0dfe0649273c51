// tpr_serve_bench -- load generator for the serving loop (tprh_server_*).
//
//   tpr_serve_bench [--rate R] [--seconds T] [--requests K] [--outstanding Q]
//                   [--clients C] [--min-log2 A] [--max-log2 B] [--max-rhs H]
//                   [--budget-gb G] [--inflight F] [--min-batch M] [--max-wait-us W]
//                   [--adaptive 0|1]
//                   [--method ptpr|tpr] [--arith fast|exact] [--seed S]
//                   [--dump FILE --dump-every D]
//
// Open loop (--rate R > 0): requests arrive as a Poisson stream of mean R per
// second for T seconds.  Closed loop (--rate 0): K requests, at most Q
// outstanding at any time (the server's saturated throughput).  Every request
// brings a matrix of n = 2^U{A..B} equations, generated on the device from
// (seed, request id) and resident in the server's arena, and 1..H right-hand
// sides generated on the device (d ~ U(-1,1)); the answers stay in the arena
// until the client sees the request complete, then request and matrix are
// released.  Latency is submit to completion on the server's clock, split
// into the wait in the queue (submit to launch) and the time on the device
// (launch to the completion the server sees).
// C client threads submit (request k belongs to client k mod C; in the closed
// loop each keeps at most Q / C outstanding) and one collector thread takes
// the completions (sleeping in tprh_server_wait_any), records latencies and
// releases; the server's calls are thread-safe.  Prints one JSON line.
// --dump: the answers of every D-th request of the schedule (k % D == 0) are
// copied out before their release and appended to FILE as records of four
// int64 (k, n, nrhs, j0) followed by nrhs * n floats, so that a test can
// check them against an independent solve; the copy runs on the collector
// thread, so a dumping run's latencies are not benchmark figures.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "tpr_hip.h"

namespace {

struct Args {
    double rate = 2000.0, seconds = 5.0;
    long long requests = 20000;
    int outstanding = 256;
    int clients = 1;
    int min_log2 = 9, max_log2 = 15, max_rhs = 64;
    double budget_gb = 64.0;
    int inflight = 4;
    int min_batch = 0;
    double max_wait_us = 0.0;
    int adaptive = 0;
    tprh_method method = TPRH_PTPR;
    tprh_arith arith = TPRH_ARITH_FAST;
    unsigned long long seed = 4044;
    std::string dump;
    long long dump_every = 0;
};

[[noreturn]] void usage(const char* msg) {
    std::fprintf(stderr, "tpr_serve_bench: %s\n", msg);
    std::exit(2);
}

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string k = argv[i];
        if (i + 1 >= argc) usage(("missing value for " + k).c_str());
        const char* v = argv[++i];
        if (k == "--rate") a.rate = std::atof(v);
        else if (k == "--seconds") a.seconds = std::atof(v);
        else if (k == "--requests") a.requests = std::atoll(v);
        else if (k == "--outstanding") a.outstanding = std::atoi(v);
        else if (k == "--clients") a.clients = std::atoi(v);
        else if (k == "--min-log2") a.min_log2 = std::atoi(v);
        else if (k == "--max-log2") a.max_log2 = std::atoi(v);
        else if (k == "--max-rhs") a.max_rhs = std::atoi(v);
        else if (k == "--budget-gb") a.budget_gb = std::atof(v);
        else if (k == "--inflight") a.inflight = std::atoi(v);
        else if (k == "--min-batch") a.min_batch = std::atoi(v);
        else if (k == "--max-wait-us") a.max_wait_us = std::atof(v);
        else if (k == "--adaptive") a.adaptive = std::atoi(v);
        else if (k == "--seed") a.seed = std::strtoull(v, nullptr, 10);
        else if (k == "--dump") a.dump = v;
        else if (k == "--dump-every") a.dump_every = std::atoll(v);
        else if (k == "--method") {
            if (!std::strcmp(v, "ptpr")) a.method = TPRH_PTPR;
            else if (!std::strcmp(v, "tpr")) a.method = TPRH_TPR;
            else usage("method must be ptpr or tpr");
        } else if (k == "--arith") {
            if (!std::strcmp(v, "fast")) a.arith = TPRH_ARITH_FAST;
            else if (!std::strcmp(v, "exact")) a.arith = TPRH_ARITH_EXACT;
            else usage("arith must be fast or exact");
        } else usage(("unknown option " + k).c_str());
    }
    if (a.min_log2 < 4 || a.max_log2 < a.min_log2 || a.max_rhs < 1) usage("bad size or right-hand-side range");
    if (a.clients < 1 || a.clients > 64) usage("clients must be 1..64");
    if (a.outstanding < a.clients) usage("outstanding must be at least the number of clients");
    if (!a.dump.empty() && a.dump_every < 1) usage("--dump needs --dump-every >= 1");
    return a;
}

void check(tprh_status st, const char* what) {
    if (st != TPRH_SUCCESS) {
        std::fprintf(stderr, "%s: %s (%s)\n", what, tprh_last_error(), tprh_status_string(st));
        std::exit(1);
    }
}

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

}  // namespace

int main(int argc, char** argv) {
    const Args a = parse(argc, argv);
    tprh_server_config cfg{};
    cfg.method = a.method;
    cfg.arith = a.arith;
    cfg.hbm_budget = (size_t)(a.budget_gb * (double)(1ull << 30));
    cfg.max_inflight = a.inflight;
    cfg.max_batch_systems = 16384;
    cfg.min_batch_systems = a.min_batch;
    cfg.max_wait_us = a.max_wait_us;
    cfg.adaptive_window = a.adaptive;
    tprh_server* srv = nullptr;
    check(tprh_server_create(&srv, &cfg), "tprh_server_create");

    // the workload: arrival times (open loop), sizes and right-hand-side counts
    std::mt19937_64 rng(a.seed);
    std::uniform_int_distribution<int> lg(a.min_log2, a.max_log2), nr(1, a.max_rhs);
    std::exponential_distribution<double> gap(a.rate > 0 ? a.rate : 1.0);
    const bool open = a.rate > 0;
    const long long nreq = open ? std::max(1LL, (long long)(a.rate * a.seconds)) : a.requests;
    std::vector<double> arrive(nreq);
    std::vector<int> n(nreq), rhs(nreq);
    double t = 0.0;
    long long systems = 0, equations = 0;
    for (long long k = 0; k < nreq; k++) {
        t += open ? gap(rng) : 0.0;
        arrive[k] = t;
        n[k] = 1 << lg(rng);
        rhs[k] = nr(rng);
        systems += rhs[k];
        equations += (long long)rhs[k] * n[k];
    }

    // warm-up: one request of every size class, served alone
    for (int e = a.min_log2; e <= a.max_log2; e++) {
        long long m, r;
        check(tprh_server_matrix_generate(srv, 1 << e, TPRH_GEN_DOMINANT, a.seed, 1ull << 40, &m), "matrix");
        check(tprh_server_submit(srv, m, 1, nullptr, a.seed, 0, nullptr, &r), "submit");
        check(tprh_server_wait(srv, r, 60.0), "wait");
        check(tprh_server_request_release(srv, r), "release");
        check(tprh_server_matrix_release(srv, m), "matrix release");
    }
    std::vector<long long> polled(4096);
    int cnt = 0;
    check(tprh_server_poll(srv, polled.data(), (int)polled.size(), &cnt), "poll");  // drop the warm-up ids
    // request ids are handed out in sequence and only this program submits:
    // the timed requests get ids first_id .. first_id + nreq - 1
    long long first_id = 0;
    {
        long long m;
        check(tprh_server_matrix_generate(srv, 1 << a.min_log2, TPRH_GEN_DOMINANT, a.seed, 1ull << 41, &m), "matrix");
        check(tprh_server_submit(srv, m, 1, nullptr, a.seed, 0, nullptr, &first_id), "submit");
        check(tprh_server_wait(srv, first_id, 60.0), "wait");
        check(tprh_server_request_release(srv, first_id), "release");
        check(tprh_server_matrix_release(srv, m), "matrix release");
        check(tprh_server_poll(srv, polled.data(), (int)polled.size(), &cnt), "poll");
        first_id++;
    }
    // owner of each request, written by its client after submit returns
    // (matrix + 1, 0 while unset) and read by the collector
    std::vector<std::atomic<long long>> owner_mat(nreq);
    std::vector<int> owner_client(nreq);
    std::vector<long long> owner_k(nreq);  // which request of the schedule
    for (auto& o : owner_mat) o.store(0);
    std::vector<std::atomic<int>> outstanding(a.clients);
    for (auto& o : outstanding) o.store(0);
    const int per_client = a.outstanding / a.clients;
    std::vector<double> lat, qwait, dev;  // submit->done, submit->launch, launch->done
    // open loop: scheduled arrival->done, so that a client that falls behind
    // its schedule (and submits late) does not hide that delay
    std::vector<double> lat_arr;
    lat_arr.reserve(nreq);
    lat.reserve(nreq);
    qwait.reserve(nreq);
    dev.reserve(nreq);
    std::atomic<long long> late_ns{0};
    std::atomic<bool> failed{false};
    const double t0 = tprh_server_now(srv);

    auto client = [&](int c) {
        for (long long k = c; k < nreq && !failed.load(); k += a.clients) {
            if (open) {
                while (tprh_server_now(srv) - t0 < arrive[k]) std::this_thread::yield();
                const long long lag = (long long)(1e9 * (tprh_server_now(srv) - t0 - arrive[k]));
                long long prev = late_ns.load();
                while (lag > prev && !late_ns.compare_exchange_weak(prev, lag)) {
                }
            } else {
                while (outstanding[c].load() >= per_client && !failed.load()) std::this_thread::yield();
            }
            long long m, r;
            check(tprh_server_matrix_generate(srv, n[k], TPRH_GEN_DOMINANT, a.seed, (uint64_t)k, &m), "matrix");
            outstanding[c]++;
            check(tprh_server_submit(srv, m, rhs[k], nullptr, a.seed, 0, nullptr, &r), "submit");
            if (r < first_id || r >= first_id + nreq) {
                std::fprintf(stderr, "unexpected request id %lld\n", r);
                std::exit(1);
            }
            owner_client[r - first_id] = c;
            owner_k[r - first_id] = k;
            owner_mat[r - first_id].store(m + 1, std::memory_order_release);
        }
    };
    FILE* dump = nullptr;
    if (!a.dump.empty() && !(dump = std::fopen(a.dump.c_str(), "wb"))) {
        std::fprintf(stderr, "cannot open %s\n", a.dump.c_str());
        return 1;
    }
    auto collector = [&]() {
        std::vector<long long> ids(4096);
        long long done = 0;
        while (done < nreq && !failed.load()) {
            int got = 0;
            check(tprh_server_wait_any(srv, ids.data(), (int)ids.size(), &got, 0.01), "wait_any");
            for (int i = 0; i < got; i++) {
                const long long r = ids[i];
                // the request may complete before its client has recorded it
                long long m1;
                while ((m1 = owner_mat[r - first_id].load(std::memory_order_acquire)) == 0) std::this_thread::yield();
                const long long matrix = m1 - 1;
                const int owner = owner_client[r - first_id];
                tprh_request_info inf;
                check(tprh_server_request_info(srv, r, &inf), "info");
                if (inf.state != 2) {
                    std::fprintf(stderr, "request %lld failed\n", r);
                    failed.store(true);
                    return;
                }
                const long long kk = owner_k[r - first_id];
                if (dump && kk % a.dump_every == 0) {
                    std::vector<float> xs((size_t)inf.nrhs * (size_t)inf.n);
                    check(tprh_server_copy_result(srv, r, xs.data()), "copy_result");
                    const long long hdr[4] = {kk, inf.n, inf.nrhs, 0};
                    if (std::fwrite(hdr, sizeof(long long), 4, dump) != 4 ||
                        std::fwrite(xs.data(), sizeof(float), xs.size(), dump) != xs.size()) {
                        std::fprintf(stderr, "writing %s failed\n", a.dump.c_str());
                        failed.store(true);
                        return;
                    }
                }
                lat.push_back(inf.t_done - inf.t_submit);
                if (open) lat_arr.push_back(inf.t_done - (t0 + arrive[owner_k[r - first_id]]));
                qwait.push_back(inf.t_launch - inf.t_submit);
                dev.push_back(inf.t_done - inf.t_launch);
                check(tprh_server_request_release(srv, r), "release");
                check(tprh_server_matrix_release(srv, matrix), "matrix release");
                outstanding[owner]--;
                done++;
            }
        }
    };
    std::thread coll(collector);
    std::vector<std::thread> cl;
    for (int c = 0; c < a.clients; c++) cl.emplace_back(client, c);
    for (auto& th : cl) th.join();
    coll.join();
    if (dump && std::fclose(dump) != 0) failed.store(true);
    if (failed.load()) return 1;
    const double late_max = 1e-9 * (double)late_ns.load();
    const double wall = tprh_server_now(srv) - t0;
    tprh_server_stats st;
    check(tprh_server_stats_get(srv, &st), "stats");
    check(tprh_server_destroy(srv), "destroy");
    // HBM traffic the solves need: 12 B per matrix equation (a, b, c) and
    // 8 B per right-hand-side equation (d read, x written)
    double mat_eq = 0.0;
    for (long long k = 0; k < nreq; k++) mat_eq += n[k];
    const double bytes = 12.0 * mat_eq + 8.0 * (double)equations;
    std::printf(
        "{\"mode\": \"%s\", \"rate\": %.1f, \"requests\": %lld, \"systems\": %lld, \"equations\": %lld, "
        "\"wall_s\": %.4f, \"requests_per_s\": %.1f, \"systems_per_s\": %.1f, \"GBps\": %.1f, "
        "\"latency_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"max\": %.4f}, "
        "\"latency_from_arrival_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"max\": %.4f}, "
        "\"queued_ms\": {\"p50\": %.4f, \"p99\": %.4f}, \"on_device_ms\": {\"p50\": %.4f, \"p99\": %.4f}, "
        "\"max_arrival_lag_ms\": %.4f, "
        "\"batches\": %lld, \"max_batch_systems_seen\": %lld, \"max_inflight_seen\": %d, \"peak_bytes\": %zu, "
        "\"outstanding\": %d, \"clients\": %d, \"min_batch\": %d, \"max_wait_us\": %.1f, \"adaptive\": %d}\n",
        open ? "poisson" : "closed", a.rate, nreq, systems, equations, wall, (double)nreq / wall,
        (double)systems / wall, bytes / wall / 1e9, 1e3 * pct(lat, 50), 1e3 * pct(lat, 99),
        1e3 * (lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end())),
        1e3 * pct(open ? lat_arr : lat, 50), 1e3 * pct(open ? lat_arr : lat, 99),
        1e3 * (open ? (lat_arr.empty() ? 0.0 : *std::max_element(lat_arr.begin(), lat_arr.end()))
                    : (lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end()))),
        1e3 * pct(qwait, 50),
        1e3 * pct(qwait, 99), 1e3 * pct(dev, 50), 1e3 * pct(dev, 99), 1e3 * late_max, st.batches,
        st.max_batch_systems_seen, st.max_inflight_seen, st.peak_bytes, open ? 0 : a.outstanding, a.clients,
        a.min_batch, a.max_wait_us, a.adaptive);
    return 0;
}
