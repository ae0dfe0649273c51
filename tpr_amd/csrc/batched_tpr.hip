// batched_tpr.hip -- the TPR instantiations of the batched kernel.
#include "batched_impl.hpp"

namespace tprh {

hipError_t launch_batched_tpr(int mode, const BatchArgs& p, int E, int L, hipStream_t stream) {
    return launch_batched_method<1>(mode, p, E, L, stream);
}

}  // namespace tprh
