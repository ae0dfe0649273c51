// capi.hip -- the extern "C" layer of libtpr_hip.so (declared in
// include/tpr_hip.h).  Host code only: argument checks, shape selection,
// device buffers of the host-pointer entry points, error reporting.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/tpr_hip.h"
#include "launch.hpp"
#include "launch_gen.hpp"

namespace {

thread_local std::string g_last_error;
thread_local bool g_timing = false;
thread_local float g_last_kernel_ms = -1.0f;

tprh_status fail(tprh_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

tprh_status fail_hip(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? TPRH_ERROR_OUT_OF_MEMORY : TPRH_ERROR_HIP;
}

#define TPRH_HIP(expr)                                      \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return fail_hip(e_, #expr);   \
    } while (0)

bool is_pow2(long long v) { return v > 0 && (v & (v - 1)) == 0; }

// Checks (n, s) for PTPR/TPR and picks the slice shape.  The reference
// requires s and m = n / s to be powers of two (src/ptpr.cpp:51-52,
// src/pcr.cpp:16) and silently misbehaves otherwise; here they are errors.
tprh_status check_ns(int n, int s, int nsys, int* E, int* L, bool explicit_shape) {
    if (n <= 0 || s <= 0 || nsys < 0) return fail(TPRH_ERROR_INVALID_VALUE, "n, s must be positive and nsys >= 0");
    if (!is_pow2(s) || s < 2) return fail(TPRH_ERROR_INVALID_VALUE, "s must be a power of two >= 2");
    if (n % s != 0) return fail(TPRH_ERROR_INVALID_VALUE, "n must be a multiple of s");
    const int m = n / s;
    if (!is_pow2(m)) return fail(TPRH_ERROR_INVALID_VALUE, "n / s must be a power of two");
    if (explicit_shape) {
        if (*E * *L != s) return fail(TPRH_ERROR_INVALID_VALUE, "E * L must equal s");
        if (!tprh::shape_compiled(*E, *L)) return fail(TPRH_ERROR_NOT_SUPPORTED, "slice shape not compiled in");
    } else if (!tprh::default_shape(s, E, L)) {
        return fail(TPRH_ERROR_NOT_SUPPORTED, "slice size s above 2048 is not supported");
    }
    if ((long long)m * *L > 1024)
        return fail(TPRH_ERROR_NOT_SUPPORTED,
                    "n / s * lanes-per-slice exceeds one workgroup (1024 lanes); use a larger s");
    return TPRH_SUCCESS;
}

tprh_status batched(tprh_method method, tprh_arith arith, const float* a, const float* b, const float* c,
                    const float* d, float* x, int n, int s, int nsys, int E, int L, bool explicit_shape,
                    void* stream) {
    if (method != TPRH_PTPR && method != TPRH_TPR)
        return fail(TPRH_ERROR_INVALID_VALUE, "batched solve: method must be TPRH_PTPR or TPRH_TPR");
    if (arith != TPRH_ARITH_FAST && arith != TPRH_ARITH_EXACT) return fail(TPRH_ERROR_INVALID_VALUE, "bad arith");
    tprh_status st = check_ns(n, s, nsys, &E, &L, explicit_shape);
    if (st != TPRH_SUCCESS) return st;
    if (nsys == 0) return TPRH_SUCCESS;
    if (!a || !c || !d || !x) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    tprh::BatchArgs p{a, b, c, d, x, n, n / s, nsys};
    hipError_t e = tprh::launch_batched(method == TPRH_PTPR ? 0 : 1, arith == TPRH_ARITH_EXACT ? 1 : 0, p, E, L,
                                        (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "batched kernel launch");
    return TPRH_SUCCESS;
}

// Device buffers for the host-pointer entry points, grown on demand and kept
// for the life of the process (the reference allocates and frees five
// buffers per call, src/ptpr.cu:554-598).
struct HostPathBuffers {
    std::mutex mu;
    float* dev = nullptr;
    size_t cap = 0;  // floats
};
HostPathBuffers g_host_bufs;

tprh_status host_path_solve(tprh_method method, tprh_arith arith, const float* a, const float* diag,
                            const float* c, const float* rhs, float* x, int n, int s) {
    if (!a || !c || !rhs || !x) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    if (n <= 0) return fail(TPRH_ERROR_INVALID_VALUE, "n must be positive");
    int E = 0, L = 0;
    if (method == TPRH_PTPR || method == TPRH_TPR) {
        tprh_status st = check_ns(n, s, 1, &E, &L, false);
        if (st != TPRH_SUCCESS) return st;
    } else if (method == TPRH_PCR || method == TPRH_CR) {
        if (!is_pow2(n)) return fail(TPRH_ERROR_INVALID_VALUE, "PCR / CR: n must be a power of two");
        if (n > tprh::kSingleMaxN) return fail(TPRH_ERROR_NOT_SUPPORTED, "PCR / CR: n above 4096 is not supported");
    } else {
        return fail(TPRH_ERROR_INVALID_VALUE, "unknown method");
    }
    std::lock_guard<std::mutex> lock(g_host_bufs.mu);
    const size_t need = 5 * (size_t)n;
    if (g_host_bufs.cap < need) {
        if (g_host_bufs.dev) (void)hipFree(g_host_bufs.dev);
        g_host_bufs.dev = nullptr;
        g_host_bufs.cap = 0;
        TPRH_HIP(hipMalloc(&g_host_bufs.dev, need * sizeof(float)));
        g_host_bufs.cap = need;
    }
    float* da = g_host_bufs.dev;
    float* dc = da + n;
    float* dr = dc + n;
    float* dx = dr + n;
    float* db = dx + n;
    const size_t bytes = sizeof(float) * (size_t)n;
    TPRH_HIP(hipMemcpy(da, a, bytes, hipMemcpyHostToDevice));
    TPRH_HIP(hipMemcpy(dc, c, bytes, hipMemcpyHostToDevice));
    TPRH_HIP(hipMemcpy(dr, rhs, bytes, hipMemcpyHostToDevice));
    if (diag) TPRH_HIP(hipMemcpy(db, diag, bytes, hipMemcpyHostToDevice));
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (g_timing) {
        TPRH_HIP(hipEventCreate(&ev[0]));
        TPRH_HIP(hipEventCreate(&ev[1]));
        TPRH_HIP(hipEventRecord(ev[0], nullptr));
    }
    hipError_t le;
    if (method == TPRH_PTPR || method == TPRH_TPR) {
        tprh::BatchArgs p{da, diag ? db : nullptr, dc, dr, dx, n, n / s, 1};
        le = tprh::launch_batched(method == TPRH_PTPR ? 0 : 1, arith == TPRH_ARITH_EXACT ? 1 : 0, p, E, L, nullptr);
    } else {
        le = tprh::launch_single(method, da, diag ? db : nullptr, dc, dr, dx, n, 1, nullptr);
    }
    g_last_kernel_ms = -1.0f;
    if (g_timing) {
        hipError_t e1 = hipEventRecord(ev[1], nullptr);
        if (le == hipSuccess && e1 == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess) {
            float ms = -1.0f;
            if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) g_last_kernel_ms = ms;
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    TPRH_HIP(le);
    TPRH_HIP(hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost));
    return TPRH_SUCCESS;
}

int ilog2i(int v) {
    int r = -1;
    while (v > 0) {
        v >>= 1;
        r++;
    }
    return r;
}

}  // namespace

struct tprh_solver {
    tprh_method method;
    tprh_arith arith;
    int n, s;
    const float *a, *diag, *c, *rhs;
    std::vector<float> x;
    bool have_system;
};

extern "C" {

const char* tprh_status_string(tprh_status status) {
    switch (status) {
        case TPRH_SUCCESS: return "success";
        case TPRH_ERROR_INVALID_VALUE: return "invalid value";
        case TPRH_ERROR_NOT_SUPPORTED: return "not supported";
        case TPRH_ERROR_HIP: return "HIP runtime error";
        case TPRH_ERROR_OUT_OF_MEMORY: return "out of memory";
        case TPRH_ERROR_BUDGET: return "memory budget exceeded";
    }
    return "unknown status";
}

const char* tprh_last_error(void) { return g_last_error.c_str(); }

int tprh_version(void) { return 100; }

tprh_status tprh_ptpr_cu(const float* a, const float* c, const float* rhs, float* x, int n, int s) {
    return host_path_solve(TPRH_PTPR, TPRH_ARITH_EXACT, a, nullptr, c, rhs, x, n, s);
}

tprh_status tprh_tpr_cu(const float* a, const float* c, const float* rhs, float* x, int n, int s) {
    return host_path_solve(TPRH_TPR, TPRH_ARITH_EXACT, a, nullptr, c, rhs, x, n, s);
}

tprh_status tprh_pcr_cu(const float* a, const float* c, const float* rhs, float* x, int n) {
    return host_path_solve(TPRH_PCR, TPRH_ARITH_EXACT, a, nullptr, c, rhs, x, n, 0);
}

tprh_status tprh_cr_cu(const float* a, const float* c, const float* rhs, float* x, int n) {
    return host_path_solve(TPRH_CR, TPRH_ARITH_EXACT, a, nullptr, c, rhs, x, n, 0);
}

tprh_status tprh_set_kernel_timing(int enable) {
    g_timing = enable != 0;
    return TPRH_SUCCESS;
}

float tprh_last_kernel_ms(void) { return g_last_kernel_ms; }

tprh_status tprh_solver_create(tprh_solver** out, tprh_method method, int n, int s) {
    if (!out) return fail(TPRH_ERROR_INVALID_VALUE, "null handle pointer");
    *out = nullptr;
    if (method == TPRH_PTPR || method == TPRH_TPR) {
        int E = 0, L = 0;
        tprh_status st = check_ns(n, s, 1, &E, &L, false);
        if (st != TPRH_SUCCESS) return st;
    } else if (method == TPRH_PCR || method == TPRH_CR) {
        if (!is_pow2(n)) return fail(TPRH_ERROR_INVALID_VALUE, "PCR / CR: n must be a power of two");
        if (n > tprh::kSingleMaxN) return fail(TPRH_ERROR_NOT_SUPPORTED, "PCR / CR: n above 4096 is not supported");
    } else {
        return fail(TPRH_ERROR_INVALID_VALUE, "unknown method");
    }
    tprh_solver* h = new (std::nothrow) tprh_solver();
    if (!h) return fail(TPRH_ERROR_OUT_OF_MEMORY, "solver allocation");
    h->method = method;
    h->arith = TPRH_ARITH_EXACT;
    h->n = n;
    h->s = s;
    h->a = h->diag = h->c = h->rhs = nullptr;
    h->x.assign((size_t)n, 0.0f);
    h->have_system = false;
    *out = h;
    return TPRH_SUCCESS;
}

tprh_status tprh_solver_set_arith(tprh_solver* h, tprh_arith arith) {
    if (!h) return fail(TPRH_ERROR_INVALID_VALUE, "null solver");
    if (arith != TPRH_ARITH_FAST && arith != TPRH_ARITH_EXACT) return fail(TPRH_ERROR_INVALID_VALUE, "bad arith");
    h->arith = arith;
    return TPRH_SUCCESS;
}

tprh_status tprh_solver_set_system(tprh_solver* h, const float* a, const float* diag, const float* c,
                                   const float* rhs) {
    if (!h) return fail(TPRH_ERROR_INVALID_VALUE, "null solver");
    if (!a || !c || !rhs) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    h->a = a;
    // CR::set_tridiagonal_system(a, diag, c, rhs) ignores diag (src/cr.cpp:17-21)
    h->diag = h->method == TPRH_CR ? nullptr : diag;
    h->c = c;
    h->rhs = rhs;
    h->have_system = true;
    return TPRH_SUCCESS;
}

tprh_status tprh_solver_solve(tprh_solver* h, int* flops) {
    if (!h) return fail(TPRH_ERROR_INVALID_VALUE, "null solver");
    if (!h->have_system) return fail(TPRH_ERROR_INVALID_VALUE, "set_tridiagonal_system was not called");
    tprh_status st = host_path_solve(h->method, h->arith, h->a, h->diag, h->c, h->rhs, h->x.data(), h->n, h->s);
    if (st != TPRH_SUCCESS) return st;
    if (flops) {
        const int n = h->n, s = h->s;
        const int m = (h->method == TPRH_PTPR || h->method == TPRH_TPR) ? n / s : 0;
        switch (h->method) {
            case TPRH_PTPR:  // src/ptpr.cpp:104-106
                *flops = m * (14 * s * ilog2i(s)) + 14 * m + 9 * (m - 1) + 14 * m * ilog2i(m) + m * 4 * (s - 1);
                break;
            case TPRH_TPR:  // src/tpr.cpp:96-98
                *flops = (28 * n - 14 * m) + (9 * (m - 1) + 17 * m) + m * 4 * (s - 1);
                break;
            case TPRH_PCR:  // src/pcr.cpp:81
                *flops = 14 * n * ilog2i(n);
                break;
            default:  // src/cr.cpp:27-31
                *flops = n > 1 ? 17 * n : 0;
                break;
        }
    }
    return TPRH_SUCCESS;
}

tprh_status tprh_solver_get_ans(tprh_solver* h, float* x) {
    if (!h) return fail(TPRH_ERROR_INVALID_VALUE, "null solver");
    if (!x) return fail(TPRH_ERROR_INVALID_VALUE, "null x");
    std::memcpy(x, h->x.data(), sizeof(float) * (size_t)h->n);
    return TPRH_SUCCESS;
}

tprh_status tprh_solver_destroy(tprh_solver* h) {
    delete h;
    return TPRH_SUCCESS;
}

tprh_status tprh_solve_batched(tprh_method method, tprh_arith arith, const float* a, const float* b, const float* c,
                               const float* d, float* x, int n, int s, int nsys, void* stream) {
    return batched(method, arith, a, b, c, d, x, n, s, nsys, 0, 0, false, stream);
}

tprh_status tprh_solve_batched_shape(tprh_method method, tprh_arith arith, const float* a, const float* b,
                                     const float* c, const float* d, float* x, int n, int s, int nsys, int E, int L,
                                     void* stream) {
    return batched(method, arith, a, b, c, d, x, n, s, nsys, E, L, true, stream);
}

int tprh_shape_supported(int E, int L) { return tprh::shape_compiled(E, L) ? 1 : 0; }

tprh_status tprh_generate(tprh_gen kind, uint64_t seed, uint64_t sys0, int nsys, int n, float* a, float* b, float* c,
                          float* d, float* xt, void* stream) {
    if (kind < TPRH_GEN_DOMINANT || kind > TPRH_GEN_RANDOM_RHS) return fail(TPRH_ERROR_INVALID_VALUE, "bad kind");
    if (n <= 0 || nsys < 0) return fail(TPRH_ERROR_INVALID_VALUE, "n must be positive and nsys >= 0");
    if (!a || !b || !c || !d) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    hipError_t e = tprh::launch_generate((int)kind, seed, sys0, nsys, n, a, b, c, d, xt, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "generate kernel launch");
    return TPRH_SUCCESS;
}

}  // extern "C"
