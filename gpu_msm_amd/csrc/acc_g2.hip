// Bucket accumulation and fix-up kernel instantiations (g2).
#include "bucket_launch.hpp"

namespace gm {

template hipError_t launch_accumulate<Fp2Ops>(hipStream_t, const Plan&, const Aff<Fp2Ops>*, uint8_t*);
template hipError_t launch_fixup<Fp2Ops>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
