// Bucket accumulation kernel instantiation (g2).
#include "engine_impl.hpp"
#include "msm_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_accumulate(hipStream_t s, const Plan& p, const Aff<O>* d_bases, uint8_t* ws) {
    const size_t nbk = (size_t)p.nw * p.B;
    k_accumulate<O><<<blocks_for((long)nbk, 64), 64, 0, s>>>(
        d_bases, reinterpret_cast<const uint32_t*>(ws + p.o_sorted), reinterpret_cast<const uint32_t*>(ws + p.o_off),
        (int)nbk, reinterpret_cast<Xyzz<O>*>(ws + p.o_buckets));
    return hipGetLastError();
}

template hipError_t launch_accumulate<Fp2Ops>(hipStream_t, const Plan&, const Aff<Fp2Ops>*, uint8_t*);

}  // namespace gm
