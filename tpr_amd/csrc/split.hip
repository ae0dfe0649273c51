// split.hip -- the split path's entry (launch_split): systems larger than
// one workgroup, dispatched to the per-method, per-arithmetic instantiations
// of split_impl.hpp.
#include <hip/hip_runtime.h>

#include "launch_split.hpp"
#include "tpr_device.hpp"

namespace tprh {

hipError_t launch_split_ptpr_fast(const SplitArgs& p, int E, int L, hipStream_t stream);
hipError_t launch_split_ptpr_exact(const SplitArgs& p, int E, int L, hipStream_t stream);
hipError_t launch_split_tpr_fast(const SplitArgs& p, int E, int L, hipStream_t stream);
hipError_t launch_split_tpr_exact(const SplitArgs& p, int E, int L, hipStream_t stream);

hipError_t launch_split(int method, int mode, const SplitArgs& p, int E, int L, hipStream_t stream) {
    if (method == 0)
        return mode == kExact ? launch_split_ptpr_exact(p, E, L, stream) : launch_split_ptpr_fast(p, E, L, stream);
    return mode == kExact ? launch_split_tpr_exact(p, E, L, stream) : launch_split_tpr_fast(p, E, L, stream);
}

}  // namespace tprh
