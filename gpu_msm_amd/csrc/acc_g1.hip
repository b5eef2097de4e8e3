// Bucket accumulation and fix-up kernel instantiations (g1).
#include "bucket_launch.hpp"

namespace gm {

template hipError_t launch_accumulate<FpOps>(hipStream_t, const Plan&, const Aff<FpOps>*, uint8_t*);
template hipError_t launch_fixup<FpOps>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
