// batched_ptpr.hip -- the PTPR instantiations of the batched kernel.
#include "batched_impl.hpp"

namespace tprh {

hipError_t launch_batched_ptpr(int mode, const BatchArgs& p, int E, int L, hipStream_t stream) {
    return launch_batched_method<0>(mode, p, E, L, stream);
}

hipError_t launch_batched(int method, int mode, const BatchArgs& p, int E, int L, hipStream_t stream) {
    return method == 0 ? launch_batched_ptpr(mode, p, E, L, stream) : launch_batched_tpr(mode, p, E, L, stream);
}

}  // namespace tprh
