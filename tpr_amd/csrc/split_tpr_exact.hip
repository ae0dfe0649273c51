// split_tpr_exact.hip -- the exact-arithmetic TPR instantiations of the split path
// (split_impl.hpp).
#include "split_impl.hpp"

namespace tprh {

hipError_t launch_split_tpr_exact(const SplitArgs& p, int E, int L, hipStream_t stream) {
    return launch_split_mode<1, kExact>(p, E, L, stream);
}

}  // namespace tprh
