// Launchers of the bucket kernels (accumulation, fix-up, bucket reduction,
// window sums), shared by the per-curve translation units. O is the storage
// policy; the kernels run with KernelOps<O> (Fp2LaneOps for G2: two lanes per
// item, so every launch has Stor<K>::LPI threads per item), the accumulation
// with AccOps<O> (for G2 the same lane pairs, single-chain column sums); a
// plan with tail_split runs the bucket reduction and window sums with
// SplitOps<O> (FpPairOps for G1, Fp2QuadOps for G2), one with fixup_split
// the fix-up's serial sums.
#pragma once
#include "engine_impl.hpp"
#include "msm_kernels.hpp"

namespace gm {

// Split lane groups of the tail kernels per storage policy.
template <class O>
struct SplitOps;
template <>
struct SplitOps<FpOps> {
    using type = FpPairOps;
};
template <>
struct SplitOps<Fp2Ops> {
    using type = Fp2QuadOps;
};

template <class O>
hipError_t launch_accumulate(hipStream_t s, const Plan& p, const Aff<O>* d_bases, uint8_t* ws) {
    using K = typename AccOps<O>::type;
    using X = typename Stor<K>::X;
    if (p.lanes <= 0) return hipSuccess;
    k_accumulate<K><<<blocks_for((long)p.lanes * Stor<K>::LPI, 64), 64, 0, s>>>(
        reinterpret_cast<const MsmDesc*>(ws + p.o_desc), p.count, (uint32_t)p.lanes, d_bases,
        reinterpret_cast<const uint64_t*>(ws + p.o_sorted), reinterpret_cast<const uint32_t*>(ws + p.o_off),
        reinterpret_cast<X*>(ws + p.o_buckets), reinterpret_cast<X*>(ws + p.o_head),
        reinterpret_cast<X*>(ws + p.o_tail));
    return hipGetLastError();
}

template <class K>
hipError_t launch_fixup_k(hipStream_t s, const Plan& p, uint8_t* ws) {
    using X = typename Stor<K>::X;
    const int nbk = (int)p.bkt_total;
    const uint32_t* off = reinterpret_cast<const uint32_t*>(ws + p.o_off);
    const X* head = reinterpret_cast<const X*>(ws + p.o_head);
    const X* tail = reinterpret_cast<const X*>(ws + p.o_tail);
    X* buckets = reinterpret_cast<X*>(ws + p.o_buckets);
    uint32_t* heavy = reinterpret_cast<uint32_t*>(ws + p.o_heavy);
    uint32_t* heavy_count = reinterpret_cast<uint32_t*>(ws + p.o_err) + 1;
    uint32_t* entries = reinterpret_cast<uint32_t*>(ws + p.o_err) + 2;
    const MsmDesc* desc = reinterpret_cast<const MsmDesc*>(ws + p.o_desc);
    k_bucket_fixup<K><<<blocks_for((long)nbk * Stor<K>::LPI, 64), 64, 0, s>>>(desc, p.count, nbk, off, head, tail,
                                                                             buckets, heavy, heavy_count, entries);
    return hipGetLastError();
}

template <class O>
hipError_t launch_fixup(hipStream_t s, const Plan& p, uint8_t* ws) {
    using K = typename KernelOps<O>::type;
    using X = typename Stor<K>::X;
    GM_TRY(p.fixup_split ? launch_fixup_k<typename SplitOps<O>::type>(s, p, ws) : launch_fixup_k<K>(s, p, ws));
    const uint32_t* off = reinterpret_cast<const uint32_t*>(ws + p.o_off);
    const X* head = reinterpret_cast<const X*>(ws + p.o_head);
    const X* tail = reinterpret_cast<const X*>(ws + p.o_tail);
    X* buckets = reinterpret_cast<X*>(ws + p.o_buckets);
    uint32_t* heavy = reinterpret_cast<uint32_t*>(ws + p.o_heavy);
    uint32_t* heavy_count = reinterpret_cast<uint32_t*>(ws + p.o_err) + 1;
    const MsmDesc* desc = reinterpret_cast<const MsmDesc*>(ws + p.o_desc);
    // one workgroup per queued heavy bucket, grid-stride over the queue (unsplit
    // either way: its threads share a bucket's partials, then a tree in LDS)
    constexpr int HT = 64 * Stor<K>::LPI;
    k_bucket_fixup_heavy<K, HT><<<2048, HT, 0, s>>>(desc, p.count, off, head, tail, buckets, heavy, heavy_count);
    return hipGetLastError();
}

template <class K>
hipError_t launch_reduce_k(hipStream_t s, const Plan& p, uint8_t* ws) {
    using X = typename Stor<K>::X;
    k_bucket_reduce<K><<<blocks_for((long)p.seg_total * Stor<K>::LPI, 64), 64, 0, s>>>(
        reinterpret_cast<const MsmDesc*>(ws + p.o_desc), p.count, p.seg_total,
        reinterpret_cast<const X*>(ws + p.o_buckets), reinterpret_cast<X*>(ws + p.o_segs));
    return hipGetLastError();
}

template <class O>
hipError_t launch_reduce(hipStream_t s, const Plan& p, uint8_t* ws) {
    if (p.seg_total == 0) return hipSuccess;
    if (p.tail_split) return launch_reduce_k<typename SplitOps<O>::type>(s, p, ws);
    return launch_reduce_k<typename KernelOps<O>::type>(s, p, ws);
}

template <class K>
hipError_t launch_window_sum_k(hipStream_t s, const Plan& p, uint8_t* ws) {
    using X = typename Stor<K>::X;
    const MsmDesc* desc = reinterpret_cast<const MsmDesc*>(ws + p.o_desc);
    // windows with few segments: one lane each (MSMs flagged DESC_WSUM_DIRECT)
    k_window_direct<K><<<blocks_for((long)p.win_total * Stor<K>::LPI, 64), 64, 0, s>>>(
        desc, p.count, p.win_total, reinterpret_cast<const X*>(ws + p.o_segs), reinterpret_cast<X*>(ws + p.o_wsum));
    GM_TRY(hipGetLastError());
    if (p.tile_total == 0) return hipSuccess;  // every window was summed directly
    k_window_partial<K><<<p.tile_total, 64 * Stor<K>::LPI, 0, s>>>(
        desc, p.count, reinterpret_cast<const X*>(ws + p.o_segs), reinterpret_cast<X*>(ws + p.o_tiles));
    GM_TRY(hipGetLastError());
    constexpr int WT = 16 * Stor<K>::LPI;
    k_window_sum<K, WT><<<p.win_total, WT, 0, s>>>(desc, p.count, reinterpret_cast<const X*>(ws + p.o_tiles),
                                                   reinterpret_cast<X*>(ws + p.o_wsum));
    return hipGetLastError();
}

template <class O>
hipError_t launch_window_sum(hipStream_t s, const Plan& p, uint8_t* ws) {
    if (p.win_total == 0) return hipSuccess;
    if (p.tail_split) return launch_window_sum_k<typename SplitOps<O>::type>(s, p, ws);
    return launch_window_sum_k<typename KernelOps<O>::type>(s, p, ws);
}

}  // namespace gm
