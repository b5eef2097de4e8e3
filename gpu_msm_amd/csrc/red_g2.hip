// Bucket-reduction and window-sum kernel instantiations (g2).
#include "bucket_launch.hpp"

namespace gm {

template hipError_t launch_reduce<Fp2Ops>(hipStream_t, const Plan&, uint8_t*);
template hipError_t launch_window_sum<Fp2Ops>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
