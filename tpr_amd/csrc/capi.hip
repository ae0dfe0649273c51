// capi.hip -- the extern "C" layer of libtpr_hip.so (declared in
// include/tpr_hip.h).  Host code only: argument checks, shape selection,
// device buffers of the host-pointer entry points, error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/tpr_hip.h"
#include "launch.hpp"
#include "launch_gen.hpp"
#include "launch_split.hpp"
#include "launch_uneven.hpp"

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

// Slice shapes measured on MI355X by the batched tuner (tpr_amd/autotune.py
// --batched, schedules/batched_gfx950.json): for PTPR and TPR, fast and
// exact, s = 256 and 512, m = n / s = 1 .. 16 slices per system and batches
// of 1, 256 and 4096 systems (2048 for n = 8192), every compiled shape with
// E * L = s that fits a workgroup was timed on inputs in HBM; a row keeps the
// fastest shape whose answers equal the default shape's on the tuning inputs
// (the default unless another is faster by more than 3 %).  Both arithmetic
// modes give the same bits with every shape (one operation order per mode),
// so the choice changes the speed only.
struct TunedShape {
    int method, arith, s, m, nsys, E, L;
};
const TunedShape kTunedShapes[] = {
#include "batched_shapes.inc"
};

int floor_log2(long long v) {
    int r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}

// The tuned shape for (method, arith, s, m): the row whose batch size is
// nearest to nsys on a log scale (the smaller on a tie); false when no row
// covers (method, arith, s, m), which keeps default_shape().
bool tuned_shape(int method, int arith, int s, int m, int nsys, int* E, int* L) {
    const int mi = method == TPRH_PTPR ? 0 : 1;
    const int ai = arith == TPRH_ARITH_EXACT ? 1 : 0;
    const int lg = floor_log2(nsys > 1 ? nsys : 1);
    const TunedShape* best = nullptr;
    int best_d = 0;
    for (const TunedShape& t : kTunedShapes) {
        if (t.method != mi || t.arith != ai || t.s != s || t.m != m) continue;
        const int d = std::abs(floor_log2(t.nsys) - lg);
        if (!best || d < best_d || (d == best_d && t.nsys < best->nsys)) {
            best = &t;
            best_d = d;
        }
    }
    if (!best) return false;
    *E = best->E;
    *L = best->L;
    return true;
}

// Checks (n, s) for PTPR/TPR and picks the slice shape.  The reference
// requires s and m = n / s to be powers of two (src/ptpr.cpp:51-52,
// src/pcr.cpp:16) and silently misbehaves otherwise; here they are errors.
// *split is set when a system needs more lanes than one workgroup has.
// wide_method >= 0: the caller's variant exists for every compiled shape
// (fp32 coefficients, no refinement), so the default may take the shape
// tuned_shape() names for that method and arithmetic.
tprh_status check_ns(int n, int s, int nsys, int* E, int* L, bool explicit_shape, bool allow_split, bool* split,
                     int wide_method = -1, int wide_arith = 0) {
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
    } else if (wide_method >= 0) {
        (void)tuned_shape(wide_method, wide_arith, s, m, nsys, E, L);
    }
    *split = (long long)m * *L > 1024;
    if (*split) {
        if (!allow_split)
            return fail(TPRH_ERROR_NOT_SUPPORTED,
                        "n / s * lanes-per-slice exceeds one workgroup (1024 lanes): use tprh_solve_ex with a "
                        "workspace, or a larger s");
        if (m > tprh::kSplitMaxM) return fail(TPRH_ERROR_NOT_SUPPORTED, "more than 2048 slices per system");
        if (!tprh::is_default_shape(*E, *L))
            return fail(TPRH_ERROR_NOT_SUPPORTED, "the multi-workgroup path uses the default slice shapes only");
    }
    return TPRH_SUCCESS;
}

// Systems per round of the multi-workgroup path: the whole batch, up to
// 64 Mi equations (a 1.75 MB workspace per million slices).
int default_chunk(int n, int nsys) {
    long long c = (64LL << 20) / n;
    if (c < 1) c = 1;
    if (c > nsys) c = nsys;
    return (int)c;
}

struct Plan {
    int E = 0, L = 0, m = 0, chunk = 0;
    bool split = false;
    size_t ws = 0;
};

tprh_status plan_desc(const tprh_solve_desc& d, Plan* pl) {
    if (d.method != TPRH_PTPR && d.method != TPRH_TPR)
        return fail(TPRH_ERROR_INVALID_VALUE, "method must be TPRH_PTPR or TPRH_TPR");
    if (d.arith != TPRH_ARITH_FAST && d.arith != TPRH_ARITH_EXACT) return fail(TPRH_ERROR_INVALID_VALUE, "bad arith");
    if (d.coef != TPRH_STORE_F32 && d.coef != TPRH_STORE_F16) return fail(TPRH_ERROR_INVALID_VALUE, "bad storage");
    if (d.refine != 0 && d.refine != 1) return fail(TPRH_ERROR_INVALID_VALUE, "refine must be 0 or 1");
    if (d.coef_shared != 0 && d.coef_shared != 1) return fail(TPRH_ERROR_INVALID_VALUE, "coef_shared must be 0 or 1");
    if (d.split_form < 0 || d.split_form > 3) return fail(TPRH_ERROR_INVALID_VALUE, "split_form must be 0 .. 3");
    if (d.x_cached != 0 && d.x_cached != 1) return fail(TPRH_ERROR_INVALID_VALUE, "x_cached must be 0 or 1");
    if (d.refine && d.coef != TPRH_STORE_F32)
        return fail(TPRH_ERROR_INVALID_VALUE, "refinement needs fp32 coefficients for the residual");
    const bool explicit_shape = d.E != 0 || d.L != 0;
    pl->E = d.E;
    pl->L = d.L;
    tprh_status st = check_ns(d.n, d.s, d.nsys, &pl->E, &pl->L, explicit_shape, true, &pl->split,
                              d.coef == TPRH_STORE_F32 && !d.refine ? (int)d.method : -1, (int)d.arith);
    if (st != TPRH_SUCCESS) return st;
    if ((d.coef == TPRH_STORE_F16 || d.refine) && !tprh::is_default_shape(pl->E, pl->L))
        return fail(TPRH_ERROR_NOT_SUPPORTED, "fp16 storage and refinement use the default slice shapes only");
    if (d.chunk < 0) return fail(TPRH_ERROR_INVALID_VALUE, "chunk must be >= 0");
    pl->m = d.n / d.s;
    pl->chunk = d.chunk > 0 ? (d.chunk < d.nsys ? d.chunk : d.nsys) : default_chunk(d.n, d.nsys);
    if (pl->chunk < 1) pl->chunk = 1;
    pl->ws = pl->split ? tprh::split_workspace_bytes(pl->m, pl->chunk) : 0;
    return TPRH_SUCCESS;
}

tprh_status solve_desc(const tprh_solve_desc& d, hipStream_t stream) {
    Plan pl;
    tprh_status st = plan_desc(d, &pl);
    if (st != TPRH_SUCCESS) return st;
    if (d.nsys == 0) return TPRH_SUCCESS;
    if (!d.a || !d.c || !d.d || !d.x) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    const int method = d.method == TPRH_PTPR ? 0 : 1;
    const int mode = d.arith == TPRH_ARITH_EXACT ? 1 : 0;
    hipError_t e;
    if (!pl.split) {
        tprh::BatchArgs p{d.a, d.b, d.c, d.d, d.x, d.n, pl.m, d.nsys};
        p.f16 = d.coef == TPRH_STORE_F16;
        p.refine = d.refine != 0;
        p.shared_coef = d.coef_shared != 0;
        p.x_cached = d.x_cached != 0;
        e = tprh::launch_batched(method, mode, p, pl.E, pl.L, stream);
    } else {
        if (!d.workspace || d.workspace_bytes < pl.ws)
            return fail(TPRH_ERROR_INVALID_VALUE, "the multi-workgroup path needs tprh_workspace_bytes() of workspace");
        tprh::SplitArgs p{};
        p.a = d.a;
        p.b = d.b;
        p.c = d.c;
        p.d = d.d;
        p.x = d.x;
        p.n = d.n;
        p.nsys = d.nsys;
        p.chunk = pl.chunk;
        p.f16 = d.coef == TPRH_STORE_F16;
        p.refine = d.refine != 0;
        p.shared_coef = d.coef_shared != 0;
        p.form = d.split_form;
        tprh::split_carve(d.workspace, pl.m, pl.chunk, &p);
        e = tprh::launch_split(method, mode, p, pl.E, pl.L, stream);
    }
    if (e != hipSuccess) return fail_hip(e, "solver kernel launch");
    return TPRH_SUCCESS;
}

tprh_status batched(tprh_method method, tprh_arith arith, const float* a, const float* b, const float* c,
                    const float* d, float* x, int n, int s, int nsys, int E, int L, bool explicit_shape,
                    void* stream) {
    if (method != TPRH_PTPR && method != TPRH_TPR)
        return fail(TPRH_ERROR_INVALID_VALUE, "batched solve: method must be TPRH_PTPR or TPRH_TPR");
    bool split = false;
    tprh_status st = check_ns(n, s, nsys, &E, &L, explicit_shape, false, &split, (int)method, (int)arith);
    if (st != TPRH_SUCCESS) return st;
    tprh_solve_desc desc{};
    desc.method = method;
    desc.arith = arith;
    desc.n = n;
    desc.s = s;
    desc.nsys = nsys;
    desc.coef = TPRH_STORE_F32;
    desc.a = a;
    desc.b = b;
    desc.c = c;
    desc.d = d;
    desc.x = x;
    desc.E = E;
    desc.L = L;
    return solve_desc(desc, (hipStream_t)stream);
}

// State of one caller of the host-pointer entry points on one device, kept
// for the life of the process (the reference allocates and frees five
// buffers per call, src/ptpr.cu:554-598): its own non-blocking stream (so a
// call never waits for other work on the device, such as torch's default
// stream), two timing events, a mapped pinned host buffer a | c | rhs | diag
// | x, and device memory for the paths that copy.  Contexts live in a pool:
// a call takes one, uses it without any lock, and gives it back, so calls
// from different threads run side by side.
struct HostCtx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t done = nullptr;  // completion of a call's work (no timing)
    float* pinned = nullptr;    // host address
    float* pinned_d = nullptr;  // the same memory as the device sees it
    size_t pcap = 0;            // floats
    float* dev = nullptr;
    size_t cap = 0;             // floats
    bool dirty = false;         // a failed call may have left work queued
};

struct HostCtxPool {
    std::mutex mu;
    std::vector<HostCtx*> idle;
};
HostCtxPool g_host_ctx;

// Takes an idle context of `device` from the pool (or a new one) and returns
// it on destruction.
struct HostCtxLease {
    HostCtx* ctx = nullptr;
    explicit HostCtxLease(int device) {
        std::lock_guard<std::mutex> g(g_host_ctx.mu);
        for (size_t i = 0; i < g_host_ctx.idle.size(); i++) {
            if (g_host_ctx.idle[i]->device == device) {
                ctx = g_host_ctx.idle[i];
                g_host_ctx.idle.erase(g_host_ctx.idle.begin() + (long)i);
                return;
            }
        }
        ctx = new (std::nothrow) HostCtx();
        if (ctx) ctx->device = device;
    }
    ~HostCtxLease() {
        if (!ctx) return;
        std::lock_guard<std::mutex> g(g_host_ctx.mu);
        g_host_ctx.idle.push_back(ctx);
    }
};

tprh_status host_ctx_ready(HostCtx& cx, size_t pinned_floats, size_t dev_floats) {
    if (!cx.stream) {
        TPRH_HIP(hipStreamCreateWithFlags(&cx.stream, hipStreamNonBlocking));
        TPRH_HIP(hipEventCreate(&cx.ev[0]));
        TPRH_HIP(hipEventCreate(&cx.ev[1]));
        TPRH_HIP(hipEventCreateWithFlags(&cx.done, hipEventDisableTiming));
    }
    if (cx.dirty) {
        TPRH_HIP(hipStreamSynchronize(cx.stream));
        cx.dirty = false;
    }
    if (cx.pcap < pinned_floats) {
        if (cx.pinned) (void)hipHostFree(cx.pinned);
        cx.pinned = cx.pinned_d = nullptr;
        cx.pcap = 0;
        TPRH_HIP(hipHostMalloc(reinterpret_cast<void**>(&cx.pinned), pinned_floats * sizeof(float),
                               hipHostMallocMapped | hipHostMallocCoherent));
        cx.pcap = pinned_floats;
        TPRH_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&cx.pinned_d), cx.pinned, 0));
    }
    if (cx.cap < dev_floats) {
        if (cx.dev) (void)hipFree(cx.dev);
        cx.dev = nullptr;
        cx.cap = 0;
        TPRH_HIP(hipMalloc(&cx.dev, dev_floats * sizeof(float)));
        cx.cap = dev_floats;
    }
    return TPRH_SUCCESS;
}

// Waits for the work a call queued on its stream.  A solve of the sizes the
// host-pointer entry points take completes within tens of microseconds, and
// a sleeping wait would add its wake-up latency to every call, so the
// completion event is polled first (for at most 2 ms), then waited for.
tprh_status host_ctx_wait(HostCtx& cx) {
    TPRH_HIP(hipEventRecord(cx.done, cx.stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t q = hipEventQuery(cx.done);
        if (q == hipSuccess) return TPRH_SUCCESS;
        if (q != hipErrorNotReady) return fail_hip(q, "hipEventQuery");
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    }
    TPRH_HIP(hipEventSynchronize(cx.done));
    return TPRH_SUCCESS;
}

// Systems up to this many equations are solved straight from the mapped
// pinned buffer: the kernel reads a, c, rhs (and diag) over the host link
// and writes x back the same way, so a call is one launch and one stream
// synchronisation, with no copy engine in between.  Larger systems (and the
// multi-workgroup path, whose kernels read the inputs twice) are copied to
// device memory first.
constexpr int kZeroCopyMaxN = 1 << 16;

tprh_status host_path_solve(tprh_method method, tprh_arith arith, const float* a, const float* diag,
                            const float* c, const float* rhs, float* x, int n, int s) {
    g_last_kernel_ms = -1.0f;
    if (!a || !c || !rhs || !x) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    if (n <= 0) return fail(TPRH_ERROR_INVALID_VALUE, "n must be positive");
    tprh_solve_desc desc{};
    Plan pl;
    if (method == TPRH_PTPR || method == TPRH_TPR) {
        desc.method = method;
        desc.arith = arith;
        desc.n = n;
        desc.s = s;
        desc.nsys = 1;
        desc.coef = TPRH_STORE_F32;
        tprh_status st = plan_desc(desc, &pl);
        if (st != TPRH_SUCCESS) return st;
    } else if (method == TPRH_PCR || method == TPRH_CR) {
        if (!is_pow2(n)) return fail(TPRH_ERROR_INVALID_VALUE, "PCR / CR: n must be a power of two");
        if (n > tprh::kSingleMaxN) return fail(TPRH_ERROR_NOT_SUPPORTED, "PCR / CR: n above 4096 is not supported");
    } else {
        return fail(TPRH_ERROR_INVALID_VALUE, "unknown method");
    }
    int device = 0;
    TPRH_HIP(hipGetDevice(&device));
    HostCtxLease lease(device);
    if (!lease.ctx) return fail(TPRH_ERROR_OUT_OF_MEMORY, "host path context");
    HostCtx& cx = *lease.ctx;
    const bool zero_copy = !pl.split && n <= kZeroCopyMaxN;
    const size_t nn = (size_t)n;
    const size_t ws_floats = (pl.ws + sizeof(float) - 1) / sizeof(float);
    tprh_status st = host_ctx_ready(cx, 5 * nn, zero_copy ? 0 : 5 * nn + ws_floats);
    if (st != TPRH_SUCCESS) return st;
    float* hp = cx.pinned;  // a | c | rhs | diag | x
    const size_t bytes = sizeof(float) * nn;
    std::memcpy(hp, a, bytes);
    std::memcpy(hp + nn, c, bytes);
    std::memcpy(hp + 2 * nn, rhs, bytes);
    if (diag) std::memcpy(hp + 3 * nn, diag, bytes);
    // kernel operands: the mapped pinned buffer itself, or device copies
    float* base = zero_copy ? cx.pinned_d : cx.dev;
    float* da = base;
    float* dc = da + nn;
    float* dr = dc + nn;
    float* db = dr + nn;
    float* dx = db + nn;
    float* ws = cx.dev ? cx.dev + 5 * nn : nullptr;
    cx.dirty = true;
    if (!zero_copy) TPRH_HIP(hipMemcpyAsync(da, hp, (diag ? 4 : 3) * bytes, hipMemcpyHostToDevice, cx.stream));
    if (g_timing) TPRH_HIP(hipEventRecord(cx.ev[0], cx.stream));
    hipError_t le = hipSuccess;
    if (method == TPRH_PTPR || method == TPRH_TPR) {
        desc.a = da;
        desc.b = diag ? db : nullptr;
        desc.c = dc;
        desc.d = dr;
        desc.x = dx;
        desc.workspace = ws;
        desc.workspace_bytes = pl.ws;
        st = solve_desc(desc, cx.stream);
    } else {
        le = tprh::launch_single(method, da, diag ? db : nullptr, dc, dr, dx, n, 1, cx.stream);
    }
    if (st != TPRH_SUCCESS) return st;
    TPRH_HIP(le);
    if (g_timing) TPRH_HIP(hipEventRecord(cx.ev[1], cx.stream));
    if (!zero_copy) TPRH_HIP(hipMemcpyAsync(hp + 4 * nn, dx, bytes, hipMemcpyDeviceToHost, cx.stream));
    st = host_ctx_wait(cx);
    if (st != TPRH_SUCCESS) return st;
    cx.dirty = false;
    if (g_timing) {
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, cx.ev[0], cx.ev[1]) == hipSuccess) g_last_kernel_ms = ms;
    }
    std::memcpy(x, hp + 4 * nn, bytes);
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

// Internal entry points for the other host-side translation units.
namespace tprh {

tprh_status set_error(tprh_status st, const std::string& msg) { return fail(st, msg); }
tprh_status set_hip_error(hipError_t e, const char* what) { return fail_hip(e, what); }

// tprh_solve_ex's plan for a caller whose workspace holds at most `cap`
// bytes, without touching the device: *chunk is the round size to solve
// with (the caller's, or with chunk 0 the default cut down until a round's
// workspace fits) and *ws_bytes what that round needs (0: no workspace).
// An explicit chunk whose round does not fit is an error.
tprh_status plan_within(const tprh_solve_desc& d, size_t cap, int* chunk, size_t* ws_bytes) {
    Plan pl;
    tprh_status st = plan_desc(d, &pl);
    if (st != TPRH_SUCCESS) return st;
    if (pl.split && pl.ws > cap) {
        const size_t one = tprh::split_workspace_bytes(pl.m, 1);
        if (d.chunk > 0 || one > cap)
            return fail(TPRH_ERROR_NOT_SUPPORTED,
                        "a round of " + std::to_string(d.chunk > 0 ? pl.chunk : 1) + " systems needs " +
                            std::to_string(d.chunk > 0 ? pl.ws : one) + " bytes of workspace; the fixed buffer holds " +
                            std::to_string(cap));
        const size_t per = one - tprh::split_workspace_bytes(pl.m, 0);
        pl.chunk = (int)((cap - tprh::split_workspace_bytes(pl.m, 0)) / per);
        pl.ws = tprh::split_workspace_bytes(pl.m, pl.chunk);
    }
    *chunk = pl.chunk;
    *ws_bytes = pl.ws;
    return TPRH_SUCCESS;
}

// Slice size and shape of the size class n under a 32-row (s, E, L) schedule
// (NULL or a zero row: s = min(n, 256), default shape).  *split: the class
// needs the multi-workgroup path.
tprh_status class_shape(int n, const int* schedule, int* s_out, int* E_out, int* L_out, bool allow_split,
                        bool* split, int method, int arith) {
    if (n < 16 || !is_pow2(n)) return fail(TPRH_ERROR_INVALID_VALUE, "system sizes must be powers of two >= 16");
    const int lg = __builtin_ctz((unsigned)n);
    int s = n < 256 ? n : 256, E = 0, L = 0;
    bool expl = false;
    if (schedule && schedule[3 * lg] != 0) {
        s = schedule[3 * lg];
        E = schedule[3 * lg + 1];
        L = schedule[3 * lg + 2];
        expl = E != 0 || L != 0;
    }
    // exact arithmetic gives the same bits with any shape, so a class without
    // an explicit shape takes the batched entry's tuned shape for one system
    // and, where that one has no path in the persistent kernels, the first
    // shape of their menu that fits
    tprh_status st = check_ns(n, s, 1, &E, &L, expl, allow_split, split,
                              arith == TPRH_ARITH_EXACT ? method : -1, arith);
    if (st != TPRH_SUCCESS) return st;
    if (arith == TPRH_ARITH_EXACT && !expl && !*split && uneven_shape_id(E, L) < 0) {
        bool found = false;
#define TPRH_MENU_TRY(ID, EE, LL)                                                 \
    if (!found && (EE) * (LL) == s && (long long)(n / s) * (LL) <= 1024) { \
        E = EE;                                                                \
        L = LL;                                                                \
        found = true;                                                          \
    }
        TPRH_UNEVEN_SHAPES(TPRH_MENU_TRY)
#undef TPRH_MENU_TRY
    }
    *s_out = s;
    *E_out = E;
    *L_out = L;
    return TPRH_SUCCESS;
}

}  // namespace tprh

struct tprh_uneven_plan {
    tprh_method method;
    tprh_arith arith;
    int nsys = 0;
    int ntasks = 0;
    int grid = 0;
    size_t lds_bytes = 0;
    long long total = 0;
    std::vector<long long> h_offsets;
    // one device allocation: tasks | sys_list | offsets | sizes | counter
    void* dev = nullptr;
    tprh::UnevenTask* tasks = nullptr;
    int* sys_list = nullptr;
    long long* offsets = nullptr;
    int* sizes = nullptr;
    int* counter = nullptr;
};

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
        case TPRH_ERROR_TIMEOUT: return "timed out";
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
        bool split = false;
        tprh_status st = check_ns(n, s, 1, &E, &L, false, true, &split);
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

tprh_status tprh_slice_shape(tprh_method method, tprh_arith arith, int n, int s, int nsys, int* E, int* L) {
    if (!E || !L) return fail(TPRH_ERROR_INVALID_VALUE, "null argument");
    if (method != TPRH_PTPR && method != TPRH_TPR) return fail(TPRH_ERROR_INVALID_VALUE, "method must be PTPR or TPR");
    if (arith != TPRH_ARITH_FAST && arith != TPRH_ARITH_EXACT) return fail(TPRH_ERROR_INVALID_VALUE, "bad arith");
    bool split = false;
    int e = 0, l = 0;
    tprh_status st = check_ns(n, s, nsys, &e, &l, false, false, &split, (int)method, (int)arith);
    if (st != TPRH_SUCCESS) return st;
    *E = e;
    *L = l;
    return TPRH_SUCCESS;
}

size_t tprh_workspace_bytes(const tprh_solve_desc* desc) {
    if (!desc) return (size_t)-1;
    Plan pl;
    if (plan_desc(*desc, &pl) != TPRH_SUCCESS) return (size_t)-1;
    return pl.ws;
}

tprh_status tprh_solve_ex(const tprh_solve_desc* desc, void* stream) {
    if (!desc) return fail(TPRH_ERROR_INVALID_VALUE, "null descriptor");
    return solve_desc(*desc, (hipStream_t)stream);
}

tprh_status tprh_uneven_plan_create(tprh_uneven_plan** out, tprh_method method, tprh_arith arith, const int* n,
                                    int nsys, const int* schedule) {
    if (!out) return fail(TPRH_ERROR_INVALID_VALUE, "null plan pointer");
    *out = nullptr;
    if (method != TPRH_PTPR && method != TPRH_TPR)
        return fail(TPRH_ERROR_INVALID_VALUE, "method must be TPRH_PTPR or TPRH_TPR");
    if (arith != TPRH_ARITH_FAST && arith != TPRH_ARITH_EXACT) return fail(TPRH_ERROR_INVALID_VALUE, "bad arith");
    if (nsys < 0 || (nsys > 0 && !n)) return fail(TPRH_ERROR_INVALID_VALUE, "bad system list");
    struct Cls {
        int s, E, L, m;
    };
    std::map<int, Cls, std::greater<int>> classes;  // largest systems first
    std::map<int, std::vector<int>> members;
    std::vector<long long> offsets((size_t)nsys + 1, 0);
    for (int k = 0; k < nsys; k++) {
        const int nk = n[k];
        if (nk < 16 || !is_pow2(nk)) return fail(TPRH_ERROR_INVALID_VALUE, "system sizes must be powers of two >= 16");
        offsets[k + 1] = offsets[k] + nk;
        if (!classes.count(nk)) {
            int s = 0, E = 0, L = 0;
            bool split = false;
            tprh_status st = tprh::class_shape(nk, schedule, &s, &E, &L, false, &split, (int)method, (int)arith);
            if (st != TPRH_SUCCESS) return st;
            if (tprh::uneven_shape_id(E, L) < 0)
                return fail(TPRH_ERROR_NOT_SUPPORTED, "slice shape has no path in the uneven-batch kernel");
            classes[nk] = Cls{s, E, L, nk / s};
        }
        members[nk].push_back(k);
    }
    std::vector<tprh::UnevenTask> tasks;
    std::vector<int> sys_list;
    size_t lds_max = 64;
    for (const auto& kv : classes) {
        const Cls& c = kv.second;
        const int T = c.m * c.L;
        const int per = 1024 / T;
        const std::vector<int>& mem = members[kv.first];
        for (size_t i = 0; i < mem.size(); i += per) {
            tprh::UnevenTask tk;
            tk.shape = tprh::uneven_shape_id(c.E, c.L);
            tk.m = c.m;
            tk.lg_lanes = __builtin_ctz((unsigned)T);
            tk.first = (int)sys_list.size();
            tk.count = (int)std::min<size_t>(per, mem.size() - i);
            for (int q = 0; q < tk.count; q++) sys_list.push_back(mem[i + q]);
            tasks.push_back(tk);
        }
        lds_max = std::max(lds_max, sizeof(float) * (size_t)per * tprh::lds_words_per_system(c.m));
    }
    tprh_uneven_plan* pl = new (std::nothrow) tprh_uneven_plan();
    if (!pl) return fail(TPRH_ERROR_OUT_OF_MEMORY, "plan allocation");
    pl->method = method;
    pl->arith = arith;
    pl->nsys = nsys;
    pl->ntasks = (int)tasks.size();
    pl->lds_bytes = lds_max;
    pl->total = offsets[nsys];
    pl->h_offsets = offsets;
    const size_t bt = sizeof(tprh::UnevenTask) * tasks.size();
    const size_t bs = sizeof(int) * sys_list.size();
    const size_t bo = sizeof(long long) * offsets.size();
    const size_t bn = sizeof(int) * (size_t)nsys;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t total = al(bt) + al(bs) + al(bo) + al(bn) + 256;
    hipError_t e = hipMalloc(&pl->dev, total);
    if (e != hipSuccess) {
        delete pl;
        return fail_hip(e, "hipMalloc(uneven plan)");
    }
    char* base = static_cast<char*>(pl->dev);
    pl->tasks = reinterpret_cast<tprh::UnevenTask*>(base);
    pl->sys_list = reinterpret_cast<int*>(base + al(bt));
    pl->offsets = reinterpret_cast<long long*>(base + al(bt) + al(bs));
    pl->sizes = reinterpret_cast<int*>(base + al(bt) + al(bs) + al(bo));
    pl->counter = reinterpret_cast<int*>(base + al(bt) + al(bs) + al(bo) + al(bn));
    std::vector<int> sizes(n, n + nsys);
    if ((bt && (e = hipMemcpy(pl->tasks, tasks.data(), bt, hipMemcpyHostToDevice)) != hipSuccess) ||
        (bs && (e = hipMemcpy(pl->sys_list, sys_list.data(), bs, hipMemcpyHostToDevice)) != hipSuccess) ||
        (e = hipMemcpy(pl->offsets, offsets.data(), bo, hipMemcpyHostToDevice)) != hipSuccess ||
        (bn && (e = hipMemcpy(pl->sizes, sizes.data(), bn, hipMemcpyHostToDevice)) != hipSuccess)) {
        (void)hipFree(pl->dev);
        delete pl;
        return fail_hip(e, "hipMemcpy(uneven plan)");
    }
    int dev = 0, cus = 0, occ = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
        (e = tprh::uneven_occupancy(method == TPRH_PTPR ? 0 : 1, arith == TPRH_ARITH_EXACT ? 1 : 0, lds_max, &occ)) !=
            hipSuccess) {
        (void)hipFree(pl->dev);
        delete pl;
        return fail_hip(e, "device query (uneven plan)");
    }
    if (occ < 1) occ = 1;
    pl->grid = std::max(1, std::min(cus * occ, pl->ntasks));
    *out = pl;
    return TPRH_SUCCESS;
}

tprh_status tprh_uneven_solve(const tprh_uneven_plan* plan, const float* a, const float* b, const float* c,
                              const float* d, float* x, void* stream) {
    if (!plan) return fail(TPRH_ERROR_INVALID_VALUE, "null plan");
    if (plan->ntasks == 0) return TPRH_SUCCESS;
    if (!a || !c || !d || !x) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    tprh::UnevenLaunch p{plan->tasks, plan->ntasks, plan->counter, plan->sys_list, plan->offsets, a, b, c, d, x,
                         plan->grid, plan->lds_bytes};
    hipError_t e = tprh::launch_uneven(plan->method == TPRH_PTPR ? 0 : 1, plan->arith == TPRH_ARITH_EXACT ? 1 : 0, p,
                                       (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "uneven kernel launch");
    return TPRH_SUCCESS;
}

long long tprh_uneven_offset(const tprh_uneven_plan* plan, int k) {
    if (!plan || k < 0 || k > plan->nsys) return -1;
    return plan->h_offsets[k];
}

long long tprh_uneven_equations(const tprh_uneven_plan* plan) { return plan ? plan->total : -1; }

tprh_status tprh_uneven_generate(const tprh_uneven_plan* plan, tprh_gen kind, uint64_t seed, uint64_t sys0, float* a,
                                 float* b, float* c, float* d, float* xt, void* stream) {
    if (!plan) return fail(TPRH_ERROR_INVALID_VALUE, "null plan");
    if (kind < TPRH_GEN_DOMINANT || kind > TPRH_GEN_RANDOM_RHS) return fail(TPRH_ERROR_INVALID_VALUE, "bad kind");
    if (!a || !b || !c || !d) return fail(TPRH_ERROR_INVALID_VALUE, "null array pointer");
    if (plan->nsys == 0) return TPRH_SUCCESS;
    hipError_t e = tprh::launch_generate_uneven((int)kind, seed, sys0, plan->nsys, plan->sizes, plan->offsets, a, b, c,
                                                d, xt, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "generate kernel launch");
    return TPRH_SUCCESS;
}

tprh_status tprh_uneven_plan_destroy(tprh_uneven_plan* plan) {
    if (plan) {
        if (plan->dev) (void)hipFree(plan->dev);
        delete plan;
    }
    return TPRH_SUCCESS;
}

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
