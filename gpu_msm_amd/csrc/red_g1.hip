// Bucket-reduction and window-sum kernel instantiations (g1).
#include "bucket_launch.hpp"

namespace gm {

template hipError_t launch_reduce<FpOps>(hipStream_t, const Plan&, uint8_t*);
template hipError_t launch_window_sum<FpOps>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
