// Bucket-reduction and window-sum kernel instantiations (g1).
#include "engine_impl.hpp"
#include "msm_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_reduce(hipStream_t s, const Plan& p, uint8_t* ws) {
    k_bucket_reduce<O><<<blocks_for((long)p.nw * p.nseg, 64), 64, 0, s>>>(
        reinterpret_cast<const Xyzz<O>*>(ws + p.o_buckets), p.nw, p.B, p.seg, reinterpret_cast<Xyzz<O>*>(ws + p.o_segs));
    return hipGetLastError();
}

template <class O>
hipError_t launch_window_sum(hipStream_t s, const Plan& p, uint8_t* ws) {
    constexpr int WT = sizeof(Xyzz<O>) > 256 ? 128 : 256;
    k_window_sum<O, WT><<<p.nw, WT, 0, s>>>(reinterpret_cast<const Xyzz<O>*>(ws + p.o_segs), p.nseg,
                                            reinterpret_cast<Xyzz<O>*>(ws + p.o_wsum));
    return hipGetLastError();
}

template hipError_t launch_reduce<FpOps>(hipStream_t, const Plan&, uint8_t*);
template hipError_t launch_window_sum<FpOps>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
