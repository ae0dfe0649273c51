// rocsparse_gtsv.cpp -- libtpr_rocsparse.so: the vendor comparator slot
// (include/tpr_rocsparse.h) over rocsparse_sgtsv_no_pivot_strided_batch.
// Host code; the library handle and the temporary buffer are kept per
// thread and grown on demand, so repeated calls of one shape allocate once.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <string>

#include "../../include/tpr_rocsparse.h"

namespace {

struct Ctx {
    rocsparse_handle h = nullptr;
    void* buf = nullptr;
    size_t cap = 0;
    ~Ctx() {
        if (buf) (void)hipFree(buf);
        if (h) (void)rocsparse_destroy_handle(h);
    }
};

thread_local Ctx g_ctx;
thread_local std::string g_err;

int fail_sp(rocsparse_status st, const char* what) {
    g_err = std::string(what) + ": rocsparse status " + std::to_string((int)st);
    return (int)st;
}

int fail_hip(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -1;
}

}  // namespace

extern "C" {

int tprh_vendor_gtsv_batch(const float* a, const float* b, const float* c, const float* d, float* x, int n, int nsys,
                           void* stream) {
    if (n < 2 || nsys < 0) {
        g_err = "n must be >= 2 and nsys >= 0";
        return (int)rocsparse_status_invalid_size;
    }
    if (nsys == 0) return 0;
    Ctx& k = g_ctx;
    rocsparse_status st;
    if (!k.h && (st = rocsparse_create_handle(&k.h)) != rocsparse_status_success)
        return fail_sp(st, "rocsparse_create_handle");
    if ((st = rocsparse_set_stream(k.h, (hipStream_t)stream)) != rocsparse_status_success)
        return fail_sp(st, "rocsparse_set_stream");
    size_t need = 0;
    if ((st = rocsparse_sgtsv_no_pivot_strided_batch_buffer_size(k.h, n, a, b, c, x, nsys, n, &need)) !=
        rocsparse_status_success)
        return fail_sp(st, "buffer_size");
    if (need > k.cap) {
        hipError_t e;
        if (k.buf && (e = hipFree(k.buf)) != hipSuccess) return fail_hip(e, "hipFree");
        k.buf = nullptr;
        k.cap = 0;
        if ((e = hipMalloc(&k.buf, need)) != hipSuccess) return fail_hip(e, "hipMalloc");
        k.cap = need;
    }
    // the solver works in place on x
    hipError_t e = hipMemcpyAsync(x, d, sizeof(float) * (size_t)n * nsys, hipMemcpyDeviceToDevice,
                                  (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpyAsync");
    if ((st = rocsparse_sgtsv_no_pivot_strided_batch(k.h, n, a, b, c, x, nsys, n, k.buf)) != rocsparse_status_success)
        return fail_sp(st, "rocsparse_sgtsv_no_pivot_strided_batch");
    return 0;
}

const char* tprh_vendor_last_error(void) { return g_err.c_str(); }

int tprh_vendor_version(void) {
    Ctx& k = g_ctx;
    if (!k.h && rocsparse_create_handle(&k.h) != rocsparse_status_success) return -1;
    int v = 0;
    if (rocsparse_get_version(k.h, &v) != rocsparse_status_success) return -1;
    return v;
}

}  // extern "C"
