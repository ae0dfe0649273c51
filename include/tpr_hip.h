/*
 * tpr_hip.h -- C ABI of the MI355X (gfx950) TPR/PTPR tridiagonal solver
 * engine, libtpr_hip.so.
 *
 * Every system is  a[i] x[i-1] + b[i] x[i] + c[i] x[i+1] = d[i],  i = 0..n-1,
 * stored as plain float arrays (structure of arrays, unit stride).  As in the
 * reference, a[0] and c[n-1] are outside the matrix and are expected to be 0;
 * n and the slice size s are powers of two with s <= n.
 *
 * Arithmetic modes:
 *   TPRH_ARITH_EXACT  the reference CPU solver's operation order, IEEE
 *                     division, no FMA: results are bit-identical to
 *                     src/ptpr.cpp / src/tpr.cpp built for x86-64 (Release).
 *   TPRH_ARITH_FAST   FMA and the hardware reciprocal in the per-equation
 *                     loops; relative L2 distance to the EXACT result stays
 *                     far below 1e-5 on diagonally dominant systems.
 *
 * Entry points and the reference interface each one replaces:
 *   tprh_ptpr_cu / tprh_tpr_cu      PTPR_CU::ptpr_cu  src/ptpr.cuh:51, src/ptpr.cu:542-600
 *                                   TPR_CU::tpr_cu    src/tpr.cuh:51,  src/tpr.cu:574-631
 *   tprh_pcr_cu / tprh_cr_cu        PTPR_CU::pcr_cu   src/ptpr.cuh:52, src/ptpr.cu:683-707
 *                                   TPR_CU::cr_cu     src/tpr.cuh:52,  src/tpr.cu:697-723
 *   tprh_solver_*                   the solver classes PTPR(n,s) / TPR(n,s):
 *                                   src/ptpr.hpp:53-97, src/tpr.hpp:53-97
 *                                   (construct, set_tridiagonal_system 3- and
 *                                   4-argument forms, solve, get_ans)
 *   tprh_solve_batched              no reference counterpart: the reference
 *                                   solves one system per launch; this solves
 *                                   a device-resident batch in one launch
 *   tprh_generate                   the seeded generator of the benchmark
 *                                   inputs (the reference's own generators,
 *                                   src/system.cpp:118-158, are unseeded)
 *
 * All functions return a tprh_status; none of them exits the process (the
 * reference's CU_CHECK exits, src/ptpr.cu:509-517).  tprh_last_error() gives
 * the message of the last failure on the calling thread.
 */
#ifndef TPR_HIP_H
#define TPR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tprh_status {
    TPRH_SUCCESS = 0,
    TPRH_ERROR_INVALID_VALUE = 1, /* bad n, s, count or null pointer */
    TPRH_ERROR_NOT_SUPPORTED = 2, /* valid, but outside what this build solves */
    TPRH_ERROR_HIP = 3,           /* a HIP runtime call failed */
    TPRH_ERROR_OUT_OF_MEMORY = 4,
    TPRH_ERROR_BUDGET = 5,        /* a serving request exceeds the memory budget */
    TPRH_ERROR_TIMEOUT = 6        /* tprh_server_wait gave up before the request completed */
} tprh_status;

typedef enum tprh_method {
    TPRH_PTPR = 0, /* PCR-like tree partitioning reduction (src/ptpr.cpp) */
    TPRH_TPR = 1,  /* tree partitioning reduction (src/tpr.cpp) */
    TPRH_PCR = 2,  /* parallel cyclic reduction (src/pcr.cpp) */
    TPRH_CR = 3    /* cyclic reduction (src/cr.cpp) */
} tprh_method;

typedef enum tprh_arith { TPRH_ARITH_FAST = 0, TPRH_ARITH_EXACT = 1 } tprh_arith;

typedef enum tprh_gen {
    TPRH_GEN_DOMINANT = 0,  /* a,c ~ U(-1,1), b = |a|+|c| + (1 + U[0,1)), xt ~ U(-1,1), d = A xt */
    TPRH_GEN_WEAK = 1,      /* as DOMINANT but b = |a|+|c| + (1e-4 + U[0,1) * 0.0099) */
    TPRH_GEN_FIXED = 2,     /* ExampleFixedInput: a = c = -1/6, b = 1, d = i + 1 */
    TPRH_GEN_RANDOM_RHS = 3 /* ExampleRandomRHSInput: a = c = -1/6, b = 1, d ~ U(-1,1) */
} tprh_gen;

const char *tprh_status_string(tprh_status status);
const char *tprh_last_error(void);
int tprh_version(void);

/* ---- drop-ins for the reference's CUDA free functions -------------------
 * Host pointers, unit diagonal, synchronous, current HIP device.
 * a, c, rhs are read only; x (n floats) receives the answer.
 * Unlike the reference, n / s > s is solved (the reduced system is not
 * limited to one block) and invalid (n, s) returns TPRH_ERROR_INVALID_VALUE
 * instead of returning silently with x untouched. */
tprh_status tprh_ptpr_cu(const float *a, const float *c, const float *rhs, float *x, int n, int s);
tprh_status tprh_tpr_cu(const float *a, const float *c, const float *rhs, float *x, int n, int s);
tprh_status tprh_pcr_cu(const float *a, const float *c, const float *rhs, float *x, int n);
tprh_status tprh_cr_cu(const float *a, const float *c, const float *rhs, float *x, int n);

/* Kernel-only timing of the host-pointer entry points above and of
 * tprh_solver_solve, as the reference's TPR_PERF build records it with device
 * events around the launch (src/ptpr.cu:574-587, src/pm.cuh:30-59): when
 * enabled, each such call records HIP events around its kernel launch and
 * tprh_last_kernel_ms() returns that duration for the calling thread's last
 * call (-1 when timing was off).  Systems of up to 65536 equations on the
 * one-workgroup path are solved straight from mapped pinned host memory (no
 * copy engine: the kernel reads a, c, rhs and writes x over the host link),
 * so for them this duration includes those transfers; larger systems are
 * copied to device memory first and the duration is the kernels' alone.
 * The host-pointer entry points are thread-safe; calls from different
 * threads run concurrently, each on a non-blocking stream of its own. */
tprh_status tprh_set_kernel_timing(int enable);
float tprh_last_kernel_ms(void);

/* ---- the solver-class surface --------------------------------------------
 * tprh_solver_create(&h, TPRH_PTPR, n, s)      ~ PTPR t(n, s);
 * tprh_solver_set_system(h, a, NULL, c, rhs)   ~ t.set_tridiagonal_system(a, c, rhs);
 * tprh_solver_set_system(h, a, diag, c, rhs)   ~ t.set_tridiagonal_system(a, diag, c, rhs);
 * tprh_solver_solve(h, &flops)                 ~ flops = t.solve();
 * tprh_solver_get_ans(h, x)                    ~ t.get_ans(x);
 * The solver keeps the host pointers given to set_system (as the reference
 * does) and reads them in solve().  The division by diag happens on the GPU;
 * the caller's arrays are not modified.  methods: TPRH_PTPR, TPRH_TPR (s used)
 * and TPRH_PCR, TPRH_CR (s ignored).  As in the reference, CR ignores diag
 * (CR::set_tridiagonal_system, src/cr.cpp:17-21). */
typedef struct tprh_solver tprh_solver;
tprh_status tprh_solver_create(tprh_solver **out, tprh_method method, int n, int s);
tprh_status tprh_solver_set_arith(tprh_solver *h, tprh_arith arith); /* default EXACT */
tprh_status tprh_solver_set_system(tprh_solver *h, const float *a, const float *diag, const float *c,
                                   const float *rhs);
tprh_status tprh_solver_solve(tprh_solver *h, int *flops);
tprh_status tprh_solver_get_ans(tprh_solver *h, float *x);
tprh_status tprh_solver_destroy(tprh_solver *h);

/* ---- device-resident batches ----------------------------------------------
 * nsys systems of n equations, back to back: system k occupies elements
 * [k*n, (k+1)*n) of a, b, c, d and x, all device pointers.  b == NULL means a
 * unit diagonal.  Asynchronous on `stream` (a hipStream_t; NULL = the null
 * stream); no host synchronisation, no allocation, graph-capturable.
 * method: TPRH_PTPR or TPRH_TPR.  The slice shape is the default for s
 * (16 lanes per slice up to s = 256, then 32 or 64) except where a tuned
 * table measured on MI355X names a faster one that gives the same answers:
 * s = 256 and 512, n / s <= 16, by method, arithmetic and batch size (the
 * tuned batch size nearest to nsys).  Answers never depend on the batch
 * size. */
tprh_status tprh_solve_batched(tprh_method method, tprh_arith arith, const float *a, const float *b,
                               const float *c, const float *d, float *x, int n, int s, int nsys,
                               void *stream);

/* Same with an explicit slice shape: E equations per lane, L lanes per slice
 * (E * L == s).  Used by the autotuner; tprh_shape_supported() lists what is
 * compiled in. */
tprh_status tprh_solve_batched_shape(tprh_method method, tprh_arith arith, const float *a, const float *b,
                                     const float *c, const float *d, float *x, int n, int s, int nsys,
                                     int E, int L, void *stream);
int tprh_shape_supported(int E, int L);
/* The slice shape (E equations per lane, L lanes per slice) that
 * tprh_solve_batched takes for this method, arithmetic, n, s and batch size;
 * no device call.  Errors as tprh_solve_batched's for (n, s). */
tprh_status tprh_slice_shape(tprh_method method, tprh_arith arith, int n, int s, int nsys, int *E, int *L);

/* ---- the general entry point ---------------------------------------------
 * Everything tprh_solve_batched does, plus:
 *   - systems larger than one workgroup (n / s * L > 1024 lanes, up to 2048
 *     slices per system): solved in rounds of `chunk` systems by three
 *     kernels (stage 1, reduced system, stage 3) through a caller-provided
 *     workspace of tprh_workspace_bytes(desc) bytes (device memory);
 *   - coefficients a, b, c stored as fp16 (TPRH_STORE_F16; d and x stay fp32;
 *     the solve computes in fp32);
 *   - one iterative-refinement step (refine = 1): x += A^-1 (d - A x) with the
 *     residual accumulated in FP64 on the GPU; x holds the current solution
 *     on entry (coefficients must be fp32).
 * No allocation, no host synchronisation, graph-capturable. */
typedef enum tprh_storage { TPRH_STORE_F32 = 0, TPRH_STORE_F16 = 1 } tprh_storage;

typedef struct tprh_solve_desc {
    tprh_method method;     /* TPRH_PTPR or TPRH_TPR */
    tprh_arith arith;
    int n, s, nsys;
    tprh_storage coef;      /* storage of a, b, c */
    const void *a, *b, *c;  /* b may be NULL: unit diagonal */
    const float *d;
    float *x;
    int refine;             /* 0: x = A^-1 d;  1: x += A^-1 (d - A x) */
    int E, L;               /* slice shape, 0 and 0 = the default for s */
    int chunk;              /* systems per round of the multi-workgroup path, 0 = default */
    void *workspace;        /* device memory, needed when a system exceeds one workgroup */
    size_t workspace_bytes;
    int coef_shared;        /* 1: every system uses the one matrix at a, b, c (n values each):
                               nsys right-hand sides of one matrix, as in ADI sweeps */
    int split_form;         /* systems above one workgroup: 0 = chosen by the library (two
                               passes); 1 = two passes (stage 1 and the reduced systems, then
                               stage 3, which reads the inputs again: 36 B/eq); 2 = one pass
                               (each slice stays in registers from load to store, 20 B/eq, where
                               every workgroup of a system can be resident at once, else two
                               passes; slower than two passes when a round has more workgroups
                               than the device holds); 3 = pipelined (one persistent launch of
                               rows, reduced-system and stage-3 tasks; slower than two passes
                               on MI355X).  Exact answers are the same bits in every form. */
    int x_cached;           /* systems within one workgroup: 0 = x written with non-temporal
                               stores (the default; measured fastest for one batch at a time and
                               for launches overlapping on several streams); 1 = x written
                               through the caches, for a caller whose next kernel reads x.
                               Same answers either way.  Larger systems: the two-pass form
                               writes x non-temporally, the one-pass and pipelined forms
                               through the caches, whatever this flag says. */
} tprh_solve_desc;

/* Workspace bytes tprh_solve_ex needs for desc (0 when none); (size_t)-1 if
 * desc is invalid. */
size_t tprh_workspace_bytes(const tprh_solve_desc *desc);
tprh_status tprh_solve_ex(const tprh_solve_desc *desc, void *stream);

/* tprh_solve_ex without any allocation.  No counterpart in the reference:
 * this is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * multi-workgroup path's workspace is a fixed device array of the library,
 * TPRH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; desc's
 * workspace must be NULL and workspace_bytes 0.  With chunk 0 the rounds are
 * cut to fit that array (one system of the largest size, 2048 slices,
 * needs 74 KB); an explicit chunk whose round needs more is refused with
 * TPRH_ERROR_NOT_SUPPORTED before anything is launched.  Calls on one GPU
 * share its array, so calls that overlap on the device race: order them on
 * one stream.  Answers are those of tprh_solve_ex, bit for bit. */
#define TPRH_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
tprh_status tprh_solve_ex_noalloc(const tprh_solve_desc *desc, void *stream);

/* ---- batches of systems of different sizes ---------------------------------
 * One persistent launch for the whole batch.  n[k] (host array) is the size
 * of system k; the systems are stored back to back in this order in a, b, c,
 * d and x.  schedule: NULL, or 32 triples (s, E, L) indexed by log2(n) giving
 * the slice size and slice shape of every size class (0, 0, 0 = default:
 * s = min(n, 256) with its default shape); a system's result depends on its
 * class's entry only, never on the rest of the batch.  Sizes: powers of two
 * from 16 with n / s * L <= 1024.  Creating the plan allocates its device
 * tables; tprh_uneven_solve allocates nothing and is graph-capturable (two
 * kernels: a counter reset and the persistent solver).  A plan holds one work
 * counter, so its solves must not overlap on the device: issue them on one
 * stream (or order them with events); solves that overlap need a plan each. */
typedef struct tprh_uneven_plan tprh_uneven_plan;
tprh_status tprh_uneven_plan_create(tprh_uneven_plan **out, tprh_method method, tprh_arith arith, const int *n,
                                    int nsys, const int *schedule);
tprh_status tprh_uneven_solve(const tprh_uneven_plan *plan, const float *a, const float *b, const float *c,
                              const float *d, float *x, void *stream);
/* Element offset of system k (host value), and the total number of equations. */
long long tprh_uneven_offset(const tprh_uneven_plan *plan, int k);
long long tprh_uneven_equations(const tprh_uneven_plan *plan);
/* The generator for the plan's layout: system k gets id sys0 + k. */
tprh_status tprh_uneven_generate(const tprh_uneven_plan *plan, tprh_gen kind, uint64_t seed, uint64_t sys0,
                                 float *a, float *b, float *c, float *d, float *xt, void *stream);
tprh_status tprh_uneven_plan_destroy(tprh_uneven_plan *plan);

/* Synthetic systems with ids sys0 .. sys0+nsys-1 written back to back into
 * device arrays a, b, c, d (and xt unless NULL); bit-identical to the
 * oracle's generator on the host.  Asynchronous on `stream`. */
tprh_status tprh_generate(tprh_gen kind, uint64_t seed, uint64_t sys0, int nsys, int n, float *a, float *b,
                          float *c, float *d, float *xt, void *stream);

/* ---- serving: a request loop with continuous batching ----------------------
 * No counterpart in the reference (its drivers solve one system per call,
 * src/pm.cpp:140-268); this is the engine a service puts in front of the
 * solver.  A server owns one device arena of `hbm_budget` bytes that holds
 * every resident matrix and every admitted request (right-hand sides,
 * answers, and the split path's workspace for systems above one workgroup).
 * Matrices stay resident until released; a request solves nrhs right-hand
 * sides against one matrix.  A worker thread admits queued requests in FIFO
 * order as arena memory frees up, groups every admitted request into one
 * persistent launch (plus one split-path launch set per size class whose n
 * exceeds a workgroup; such classes need s >= 256), and keeps up to
 * `max_inflight` batches on the device at once,
 * one stream each, so requests that arrive while others run are launched
 * beside them.  A system's answer depends only on its matrix, right-hand
 * side and size class (n -> s, E, L), never on the batch it rides in.
 * All entry points are thread-safe. */
typedef struct tprh_server tprh_server;

typedef struct tprh_server_config {
    tprh_method method;     /* TPRH_PTPR or TPRH_TPR */
    tprh_arith arith;
    size_t hbm_budget;      /* bytes of the arena (matrices + admitted requests) */
    int max_inflight;       /* batches on the device at once (>= 1) */
    int max_batch_systems;  /* right-hand sides per batch (>= 1) */
    const int *schedule;    /* NULL or 32 (s, E, L) triples as tprh_uneven_plan_create */
    /* Batching window.  With nothing on the device a batch is launched at
     * once; while batches run, the next one waits until min_batch_systems
     * right-hand sides are queued or the oldest queued request has waited
     * max_wait_us (0, 0: launch whenever a slot is free). */
    int min_batch_systems;
    double max_wait_us;
    /* 1: adaptive window instead of the fixed one.  The next batch waits
     * (a) while requests arrive faster than one per launch time (the time
     * the server takes to build and launch a batch, measured), until it
     * holds about twice the requests that arrive in one launch time -- so
     * that under heavy load batches stay large and the worker is not the
     * bottleneck; and (b) while batches run, until the earliest of them is
     * expected to finish within one launch time (device time per equation
     * is measured as batches complete).  It never waits past
     * max_batch_systems right-hand sides, nor past max_wait_us for its
     * oldest request (four launch times when max_wait_us is 0);
     * min_batch_systems is not used.  Under light load nothing waits.  0:
     * the fixed window above. */
    int adaptive_window;
} tprh_server_config;

tprh_status tprh_server_create(tprh_server **out, const tprh_server_config *cfg);
/* Waits for every queued and in-flight request, then frees everything. */
tprh_status tprh_server_destroy(tprh_server *srv);

/* Resident matrices of n equations (n a power of two >= 16); fails with
 * TPRH_ERROR_BUDGET when the arena has no room.  _upload copies a, b, c
 * (host or device pointers; the copy is complete when the call returns);
 * _generate makes no device call: the first batch that solves a request on
 * the matrix fills it on the device with the seeded generator (system id
 * `sys`, kinds as tprh_generate), in the same launch as the batch's
 * synthetic right-hand sides. */
tprh_status tprh_server_matrix_upload(tprh_server *srv, int n, const float *a, const float *b, const float *c,
                                      long long *matrix);
tprh_status tprh_server_matrix_generate(tprh_server *srv, int n, tprh_gen kind, uint64_t seed, uint64_t sys,
                                        long long *matrix);
/* Freed once no queued or in-flight request uses it. */
tprh_status tprh_server_matrix_release(tprh_server *srv, long long matrix);

/* Queues nrhs right-hand sides against `matrix`.  d: nrhs * n floats (host
 * or device; copied when the request is launched, so it must stay valid
 * until then), or NULL to generate right-hand side j as
 * d ~ U(-1,1) of (rhs_seed, the matrix's generator id, j0 + j) (id 0 for an
 * uploaded matrix).  x: where
 * the answers go (nrhs * n floats, host or device) when the request
 * completes, or NULL to keep them in the arena until
 * tprh_server_request_release (tprh_server_request_info gives the device
 * address).  TPRH_ERROR_BUDGET: the request needs more than the whole arena.
 * Host d and x should be pinned (hipHostMalloc) or device memory: both are
 * copied by hipMemcpyAsync on the batch's stream (d before its kernels, x
 * after them), and a copy from or to pageable host memory may make the
 * server's worker thread wait for that batch, so that batches no longer
 * overlap.  Pageable buffers give the same answers. */
tprh_status tprh_server_submit(tprh_server *srv, long long matrix, int nrhs, const float *d, uint64_t rhs_seed,
                               int j0, float *x, long long *request);

typedef struct tprh_request_info {
    int state;           /* 0 queued, 1 on the device, 2 done, 3 failed */
    int n, nrhs;
    double t_submit;     /* seconds on the server's clock (tprh_server_now) */
    double t_launch;
    double t_done;       /* when the worker saw the batch complete */
    const float *x;      /* device address of the answers while retained, else NULL */
} tprh_request_info;

tprh_status tprh_server_request_info(tprh_server *srv, long long request, tprh_request_info *info);
/* Blocks until the request is done (TPRH_SUCCESS), failed (its error), or
 * timeout_s passes (TPRH_ERROR_TIMEOUT; the request keeps going).  timeout_s
 * must be >= 0 (INFINITY is allowed; waits beyond a year end after a year). */
tprh_status tprh_server_wait(tprh_server *srv, long long request, double timeout_s);
/* Ids of requests completed since the previous call (at most max), oldest
 * first; requests already released are left out. */
tprh_status tprh_server_poll(tprh_server *srv, long long *ids, int max, int *count);
/* As tprh_server_poll, but sleeps until at least one request has completed
 * or timeout_s passes (then *count is 0).  Lets a client wait for whichever
 * request finishes first without spinning on the server. */
tprh_status tprh_server_wait_any(tprh_server *srv, long long *ids, int max, int *count, double timeout_s);
/* Copies a completed request's retained answers (nrhs * n floats) to dst
 * (host or device); returns when the copy is done. */
tprh_status tprh_server_copy_result(tprh_server *srv, long long request, float *dst);
/* Forgets a completed request and frees its retained answers. */
tprh_status tprh_server_request_release(tprh_server *srv, long long request);
double tprh_server_now(const tprh_server *srv);

typedef struct tprh_server_stats {
    long long submitted, completed, failed;
    long long systems, equations;  /* right-hand sides and equations solved */
    long long batches, max_batch_systems_seen;
    int max_inflight_seen;
    int queued;
    size_t bytes_in_use, peak_bytes, budget;
} tprh_server_stats;

tprh_status tprh_server_stats_get(tprh_server *srv, tprh_server_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* TPR_HIP_H */
