// noalloc.hip -- tprh_solve_ex_noalloc: tprh_solve_ex for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The multi-workgroup path's workspace
// (per slice: the boundary rows and x at the slice end) comes from a fixed
// device array of this library instead of a caller's buffer.
//
// `scratch` below is a device symbol: in host code its name is only a
// stand-in, so the host never passes it to a kernel or a copy.  Its device
// address, different on every GPU, is looked up once per device with
// hipGetSymbolAddress (with that device current) on the first call there,
// kept in g_scratch_addr, and that address is what the kernels receive as
// their workspace.
//
// Every call on one GPU shares that GPU's array: calls that overlap on the
// device (different streams without ordering between them) race on it.
// Nothing in it needs clearing by the caller: each round of the split path
// writes every word it later reads (the one-pass form's tickets and flags are
// zeroed by a fill kernel at the start of its round, not by hipMemsetAsync).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/tpr_hip.h"
#include "launch_split.hpp"

namespace tprh {
tprh_status set_error(tprh_status st, const std::string& msg);
tprh_status set_hip_error(hipError_t e, const char* what);
tprh_status plan_within(const tprh_solve_desc& d, size_t cap, int* chunk, size_t* ws_bytes);
}  // namespace tprh

namespace {

constexpr size_t kScratchBytes = TPRH_NOALLOC_SCRATCH_BYTES;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
tprh_status scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return tprh::set_hip_error(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return tprh::set_error(TPRH_ERROR_NOT_SUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return tprh::set_hip_error(e, "hipGetSymbolAddress(scratch)");
        if (!p) return tprh::set_error(TPRH_ERROR_HIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return TPRH_SUCCESS;
}

}  // namespace

extern "C" tprh_status tprh_solve_ex_noalloc(const tprh_solve_desc* desc, void* stream) {
    if (!desc) return tprh::set_error(TPRH_ERROR_INVALID_VALUE, "null descriptor");
    if (desc->workspace || desc->workspace_bytes)
        return tprh::set_error(TPRH_ERROR_INVALID_VALUE,
                               "tprh_solve_ex_noalloc takes no workspace: leave workspace NULL and workspace_bytes 0");
    tprh_solve_desc d = *desc;
    int chunk = 0;
    size_t need = 0;
    tprh_status st = tprh::plan_within(d, kScratchBytes, &chunk, &need);
    if (st != TPRH_SUCCESS) return st;
    d.chunk = chunk;
    if (need > 0) {
        void* p = nullptr;
        st = scratch_address(&p);
        if (st != TPRH_SUCCESS) return st;
        d.workspace = p;
        d.workspace_bytes = kScratchBytes;
    }
    return tprh_solve_ex(&d, stream);
}
