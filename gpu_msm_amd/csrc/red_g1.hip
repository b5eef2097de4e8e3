// Bucket-reduction and window-sum kernel instantiations (g1).
#include "engine_impl.hpp"
#include "msm_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_reduce(hipStream_t s, const Plan& p, uint8_t* ws) {
    if (p.seg_total == 0) return hipSuccess;
    k_bucket_reduce<O><<<blocks_for((long)p.seg_total, 64), 64, 0, s>>>(
        reinterpret_cast<const MsmDesc*>(ws + p.o_desc), p.count, p.seg_total,
        reinterpret_cast<const Xyzz<O>*>(ws + p.o_buckets), reinterpret_cast<Xyzz<O>*>(ws + p.o_segs));
    return hipGetLastError();
}

template <class O>
hipError_t launch_window_sum(hipStream_t s, const Plan& p, uint8_t* ws) {
    if (p.win_total == 0) return hipSuccess;
    const MsmDesc* desc = reinterpret_cast<const MsmDesc*>(ws + p.o_desc);
    k_window_partial<O><<<p.tile_total, 64, 0, s>>>(desc, p.count, reinterpret_cast<const Xyzz<O>*>(ws + p.o_segs),
                                                    reinterpret_cast<Xyzz<O>*>(ws + p.o_tiles));
    GM_TRY(hipGetLastError());
    k_window_sum<O, 16><<<p.win_total, 16, 0, s>>>(desc, p.count, reinterpret_cast<const Xyzz<O>*>(ws + p.o_tiles),
                                                   reinterpret_cast<Xyzz<O>*>(ws + p.o_wsum));
    return hipGetLastError();
}

template hipError_t launch_reduce<FpOps>(hipStream_t, const Plan&, uint8_t*);
template hipError_t launch_window_sum<FpOps>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
