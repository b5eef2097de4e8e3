// Bucket accumulation and fix-up kernel instantiations (g2).
#include "engine_impl.hpp"
#include "msm_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_accumulate(hipStream_t s, const Plan& p, const Aff<O>* d_bases, uint8_t* ws) {
    const int nbk = (int)p.bkt_total;
    const uint32_t* off = reinterpret_cast<const uint32_t*>(ws + p.o_off);
    Xyzz<O>* buckets = reinterpret_cast<Xyzz<O>*>(ws + p.o_buckets);
    Xyzz<O>* head = reinterpret_cast<Xyzz<O>*>(ws + p.o_head);
    Xyzz<O>* tail = reinterpret_cast<Xyzz<O>*>(ws + p.o_tail);
    (void)nbk;
    if (p.lanes > 0) {
        k_accumulate<O><<<blocks_for(p.lanes, 64), 64, 0, s>>>(
            reinterpret_cast<const MsmDesc*>(ws + p.o_desc), p.count, (uint32_t)p.lanes, d_bases,
            reinterpret_cast<const uint64_t*>(ws + p.o_sorted), off, buckets, head, tail);
        GM_TRY(hipGetLastError());
    }
    return hipSuccess;
}

template <class O>
hipError_t launch_fixup(hipStream_t s, const Plan& p, uint8_t* ws) {
    const int nbk = (int)p.bkt_total;
    const uint32_t* off = reinterpret_cast<const uint32_t*>(ws + p.o_off);
    const Xyzz<O>* head = reinterpret_cast<const Xyzz<O>*>(ws + p.o_head);
    const Xyzz<O>* tail = reinterpret_cast<const Xyzz<O>*>(ws + p.o_tail);
    Xyzz<O>* buckets = reinterpret_cast<Xyzz<O>*>(ws + p.o_buckets);
    uint32_t* heavy = reinterpret_cast<uint32_t*>(ws + p.o_heavy);
    uint32_t* heavy_count = reinterpret_cast<uint32_t*>(ws + p.o_err) + 1;
    const MsmDesc* desc = reinterpret_cast<const MsmDesc*>(ws + p.o_desc);
    k_bucket_fixup<O><<<blocks_for(nbk, 64), 64, 0, s>>>(desc, p.count, nbk, off, head, tail, buckets, heavy,
                                                         heavy_count);
    GM_TRY(hipGetLastError());
    // one wave per queued heavy bucket, grid-stride over the queue
    k_bucket_fixup_heavy<O, 64><<<2048, 64, 0, s>>>(desc, p.count, off, head, tail, buckets, heavy, heavy_count);
    return hipGetLastError();
}

template hipError_t launch_accumulate<Fp2Ops>(hipStream_t, const Plan&, const Aff<Fp2Ops>*, uint8_t*);
template hipError_t launch_fixup<Fp2Ops>(hipStream_t, const Plan&, uint8_t*);

}  // namespace gm
