// Pippenger MSM kernels for gfx950 (G1 and G2 share every kernel through the
// field policy O).
//
// Pipeline for one MSM over N points with c-bit signed windows,
// nw windows and B = 2^(c-1) buckets per window:
//   k_digits       signed-digit recoding of each scalar into 16-bit codes
//                  (reference src/curve.cu:892-922 extracts unsigned windows;
//                  here digits are in [-2^(c-1), 2^(c-1)] and the sign negates
//                  the base's y)
//   k_tile_count / scan / k_tile_place  counting sort of the (window, point)
//                  entries by bucket, tiles counted in LDS (sort_kernels.hpp)
//   k_accumulate   each lane mixed-adds a fixed chunk of sorted entries into
//                  XYZZ accumulators; runs cut by a chunk edge go to head/tail
//                  partials that k_bucket_fixup sums
//   k_bucket_reduce one lane per segment of buckets: running sums give
//                  sum_b (b+1) B_b for the segment as T + lo * S
//   k_window_partial / k_window_sum  tiles of segments, then each window's
//                  total over its tiles (two short tree stages); windows of
//                  few segments: k_window_direct, one lane per window
//   combine        Horner over windows, most significant first (reference
//                  src/curve.cu:2060-2179 combines on the CPU after every
//                  window; here on the host once, or k_horner on the device)
#pragma once
#include "curve.hpp"
#include "desc_find.hpp"
#include "pair_curve.hpp"

namespace gm {

template <class T>
__device__ __forceinline__ void ld16(T& dst, const T* src) {
    static_assert(sizeof(T) % 16 == 0, "16-byte granularity");
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(&dst);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); ++i) d[i] = s[i];
}
template <class T>
__device__ __forceinline__ void st16(T* dst, const T& src) {
    static_assert(sizeof(T) % 16 == 0, "16-byte granularity");
    uint4* d = reinterpret_cast<uint4*>(dst);
    const uint4* s = reinterpret_cast<const uint4*>(&src);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); ++i) d[i] = s[i];
}

// ------------------------------------------------------------ format import

// 7 x u64 (any 448-bit value) -> 16 x 28-bit limbs, reduced mod p.
__device__ __forceinline__ void fp_from_u64x7(Fp& r, const uint64_t* w) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int bit = 28 * i;
        uint64_t x = w[bit / 64] >> (bit % 64);
        if (bit % 64 > 36) x |= w[bit / 64 + 1] << (64 - bit % 64);
        r.v[i] = (uint32_t)x & MASK28;
    }
    // 2^448 < 5p: four conditional subtractions reach [0, p)
#pragma unroll
    for (int k = 0; k < 4; ++k) fp_canon(r, r);
}

__device__ __forceinline__ void fp_to_u64x7(uint64_t* w, const Fp& a) {
#pragma unroll
    for (int j = 0; j < 7; ++j) w[j] = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int bit = 28 * i;
        const uint64_t x = a.v[i];
        w[bit / 64] |= x << (bit % 64);
        if (bit % 64 > 36) w[bit / 64 + 1] |= x >> (64 - bit % 64);
    }
}

// Every element is stored canonical, in [0, p): the accumulation negates a
// base's y as p - y. Montgomery-form input (below 2^448 < 5p) is reduced by up
// to four subtractions of p.
__device__ __forceinline__ void import_elem(Fp& r, const uint64_t* w, bool to_mont) {
    fp_from_u64x7(r, w);
    if (to_mont) {
        Fp r2;
        fp_set_const(r2, R2L);
        fp_mul(r, r, r2);
        fp_canon(r, r);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) fp_cond_sub(r.v, PL);
    }
}
__device__ __forceinline__ void import_elem(Fp2& r, const uint64_t* w, bool to_mont) {
    import_elem(r.c0, w, to_mont);
    import_elem(r.c1, w + 7, to_mont);
}
__device__ __forceinline__ void export_elem(uint64_t* w, const Fp& a) {
    Fp c;
    fp_canon(c, a);
    fp_to_u64x7(w, c);
}
__device__ __forceinline__ void export_elem(uint64_t* w, const Fp2& a) {
    export_elem(w, a.c0);
    export_elem(w + 7, a.c1);
}

template <class O>
__device__ __forceinline__ void aff_set_inf(Aff<O>& p) {
    O::zero(p.x);
    O::zero(p.y);
    uint32_t* xw = reinterpret_cast<uint32_t*>(&p.x);
    xw[NL - 1] = INF_MARK;
    if (sizeof(p.x) == 2 * sizeof(Fp)) xw[2 * NL - 1] = INF_MARK;
}

// ---------------------------------------------- storage of the hot kernels
// Bases, buckets and partial sums are stored as full Aff / Xyzz records; the
// G2 bucket kernels compute with Fp2LaneOps, two lanes per record, each lane
// loading and storing its own component (Stor<O>::LPI lanes per item).
template <class O>
struct Stor {
    using A = Aff<O>;
    using X = Xyzz<O>;
    static constexpr int LPI = 1;
};
template <>
struct Stor<Fp2LaneOps> {
    using A = Aff<Fp2Ops>;
    using X = Xyzz<Fp2Ops>;
    static constexpr int LPI = 2;
};
template <>
struct Stor<Fp2LaneAccOps> : Stor<Fp2LaneOps> {};
// The G1 accumulation's single-chain policy stores plain G1 records.
template <>
struct Stor<FpAccOps> {
    using A = Aff<FpOps>;
    using X = Xyzz<FpOps>;
    static constexpr int LPI = 1;
};
// Tail kernels on split lane groups: G1 two lanes per item, each holding the
// whole record; G2 four (two lane pairs), each pair holding the record.
template <>
struct Stor<FpPairOps> {
    using A = Aff<FpOps>;
    using X = Xyzz<FpOps>;
    static constexpr int LPI = 2;
};
template <>
struct Stor<Fp2QuadOps> {
    using A = Aff<Fp2Ops>;
    using X = Xyzz<Fp2Ops>;
    static constexpr int LPI = 4;
};

template <class O>
__device__ __forceinline__ void ldp(Aff<O>& d, const Aff<O>* s) {
    ld16(d, s);
}
template <class O>
__device__ __forceinline__ void ldp(Xyzz<O>& d, const Xyzz<O>* s) {
    ld16(d, s);
}
template <class O>
__device__ __forceinline__ void stp(Xyzz<O>* d, const Xyzz<O>& v) {
    st16(d, v);
}
template <bool C>
__device__ __forceinline__ void ldp(Aff<Fp2LaneOpsT<C>>& d, const Aff<Fp2Ops>* s) {
    const uint32_t q = lane_part();
    ld16(d.x, &s->x.c0 + q);
    ld16(d.y, &s->y.c0 + q);
}
template <bool C>
__device__ __forceinline__ void ldp(Xyzz<Fp2LaneOpsT<C>>& d, const Xyzz<Fp2Ops>* s) {
    const uint32_t q = lane_part();
    ld16(d.X, &s->X.c0 + q);
    ld16(d.Y, &s->Y.c0 + q);
    ld16(d.ZZ, &s->ZZ.c0 + q);
    ld16(d.ZZZ, &s->ZZZ.c0 + q);
}
__device__ __forceinline__ void ldp(Aff<FpAccOps>& d, const Aff<FpOps>* s) {
    ld16(d, reinterpret_cast<const Aff<FpAccOps>*>(s));
}
__device__ __forceinline__ void stp(Xyzz<FpOps>* d, const Xyzz<FpAccOps>& v) {
    st16(reinterpret_cast<Xyzz<FpAccOps>*>(d), v);
}
__device__ __forceinline__ void ldp(Xyzz<FpPairOps>& d, const Xyzz<FpOps>* s) {
    ld16(d, reinterpret_cast<const Xyzz<FpPairOps>*>(s));
}
// Both lanes of a pair hold the same record; the first stores it.
__device__ __forceinline__ void stp(Xyzz<FpOps>* d, const Xyzz<FpPairOps>& v) {
    if (lane_part() == 0) st16(reinterpret_cast<Xyzz<FpPairOps>*>(d), v);
}
__device__ __forceinline__ void ldp(Xyzz<Fp2QuadOps>& d, const Xyzz<Fp2Ops>* s) {
    const uint32_t q = lane_part();
    ld16(d.X, &s->X.c0 + q);
    ld16(d.Y, &s->Y.c0 + q);
    ld16(d.ZZ, &s->ZZ.c0 + q);
    ld16(d.ZZZ, &s->ZZZ.c0 + q);
}
// Both pairs of a quad hold the same record; the first stores it.
__device__ __forceinline__ void stp(Xyzz<Fp2Ops>* d, const Xyzz<Fp2QuadOps>& v) {
    if (quad_half() != 0) return;
    const uint32_t q = lane_part();
    st16(&d->X.c0 + q, v.X);
    st16(&d->Y.c0 + q, v.Y);
    st16(&d->ZZ.c0 + q, v.ZZ);
    st16(&d->ZZZ.c0 + q, v.ZZZ);
}
template <bool C>
__device__ __forceinline__ void stp(Xyzz<Fp2Ops>* d, const Xyzz<Fp2LaneOpsT<C>>& v) {
    const uint32_t q = lane_part();
    st16(&d->X.c0 + q, v.X);
    st16(&d->Y.c0 + q, v.Y);
    st16(&d->ZZ.c0 + q, v.ZZ);
    st16(&d->ZZZ.c0 + q, v.ZZZ);
}

// Compute policy of the bucket kernels for a curve's storage policy.
template <class O>
struct KernelOps {
    using type = O;
};
template <>
struct KernelOps<Fp2Ops> {
    using type = Fp2LaneOps;
};
// Policy of the accumulation kernel alone: for G2 the lane pairs with
// single-chain column sums (two waves per SIMD hide the longer dependency
// chain there; the fix-up and tail kernels keep Fp2LaneOps).
template <class O>
struct AccOps {
    using type = typename KernelOps<O>::type;
};
template <>
struct AccOps<Fp2Ops> {
    using type = Fp2LaneAccOps;
};
template <>
struct AccOps<FpOps> {
    using type = FpAccOps;
};

// ABI affine records (reference include/curve.h:13-25) -> device affine.
template <class O>
__global__ void k_import_aff(Aff<O>* __restrict__ out, const uint8_t* __restrict__ in, int n, int to_mont) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using L = AbiLayout<O>;
    const uint8_t* rec = in + (size_t)i * L::AFF_BYTES;
    const uint64_t* w = reinterpret_cast<const uint64_t*>(rec);
    Aff<O> p;
    if (rec[2 * L::ELEM_U64 * 8] != 0) {
        aff_set_inf(p);
    } else {
        import_elem(p.x, w, to_mont != 0);
        import_elem(p.y, w + L::ELEM_U64, to_mont != 0);
    }
    st16(&out[i], p);
}

// device affine -> ABI affine records, Montgomery form.
template <class O>
__global__ void k_export_aff(uint8_t* __restrict__ out, const Aff<O>* __restrict__ in, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using L = AbiLayout<O>;
    Aff<O> p;
    ld16(p, &in[i]);
    uint64_t w[2 * L::ELEM_U64 + 1];
#pragma unroll
    for (int j = 0; j < 2 * L::ELEM_U64 + 1; ++j) w[j] = 0;
    if (aff_is_inf(p)) {
        w[2 * L::ELEM_U64] = 1;  // bool infinity = true, padding zero
    } else {
        export_elem(w, p.x);
        export_elem(w + L::ELEM_U64, p.y);
    }
    uint64_t* o = reinterpret_cast<uint64_t*>(out + (size_t)i * L::AFF_BYTES);
#pragma unroll
    for (int j = 0; j < 2 * L::ELEM_U64 + 1; ++j) o[j] = w[j];
}

// Bases P_i = k_i * G for 64-bit k_i (double-and-add, then affine).
template <class O>
__global__ void __launch_bounds__(64) k_gen_bases(Aff<O>* __restrict__ out, const uint64_t* __restrict__ k, int n,
                                                  Aff<O> g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t s = k[i];
    Aff<O> r;
    if (s == 0) {
        aff_set_inf(r);
    } else {
        Xyzz<O> acc;
        xyzz_set_inf(acc);
        const int top = 63 - __builtin_clzll(s);
        for (int b = top; b >= 0; --b) {
            xyzz_dbl<O>(acc, acc);
            if ((s >> b) & 1) xyzz_add_aff<O>(acc, g.x, g.y);
        }
        if (!xyzz_to_aff<O>(r.x, r.y, acc)) aff_set_inf(r);
    }
    st16(&out[i], r);
}

// ------------------------------------------------------ bucket accumulation

// Entries are 64-bit (global bucket << 32) | point index | sign << 31, sorted
// so each bucket's entries are contiguous (offsets in off[]) and each MSM's
// buckets follow each other. Every lane takes a fixed chunk of its MSM's
// entries (desc.chunk of them), whatever the bucket sizes, so lanes stay
// balanced even when a window has few, very large buckets (the top window of
// a 299-bit scalar holds only a few bits). A run of one bucket wholly inside a
// chunk is written to its bucket; a run cut by the chunk's start goes to
// head[lane], one cut only by its end to tail[lane], and k_bucket_fixup sums
// those partials.
template <class O>
__global__ void __launch_bounds__(64, 2) k_accumulate(const MsmDesc* __restrict__ desc, int count, uint32_t lanes,
                                                   const typename Stor<O>::A* __restrict__ bases,
                                                   const uint64_t* __restrict__ entries,
                                                   const uint32_t* __restrict__ off, typename Stor<O>::X* __restrict__ buckets,
                                                   typename Stor<O>::X* __restrict__ head, typename Stor<O>::X* __restrict__ tail) {
    const uint32_t l = (blockIdx.x * blockDim.x + threadIdx.x) / Stor<O>::LPI;
    if (l >= lanes) return;
    const MsmDesc& m = desc[GM_FIND(desc, count, lane_base, l)];
    const uint32_t e0 = off[m.bkt_base], e1 = off[m.bkt_base + m.nw * m.B];
    const uint32_t j0 = e0 + (l - m.lane_base) * m.chunk;
    if (j0 >= e1) return;
    const uint32_t j1 = min(j0 + m.chunk, e1);
    Xyzz<O> acc;
    xyzz_set_inf(acc);
    uint32_t cur = (uint32_t)(entries[j0] >> 32);
    bool first = true;
    // G1 (one lane per item, 202 registers) loads the next entry's base while
    // the current mixed add runs; a G2 lane pair has no registers to spare.
    constexpr bool PREFETCH = Stor<O>::LPI == 1;
    uint64_t en = entries[j0];
    Aff<O> nx;
    if constexpr (PREFETCH) ldp(nx, &bases[(uint32_t)en & 0x7FFFFFFFu]);
    for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t e = en;
        Aff<O> pt;
        if constexpr (PREFETCH) {
            pt = nx;
            if (j + 1 < j1) {
                en = entries[j + 1];
                ldp(nx, &bases[(uint32_t)en & 0x7FFFFFFFu]);
            }
        } else {
            ldp(pt, &bases[(uint32_t)e & 0x7FFFFFFFu]);
            if (j + 1 < j1) en = entries[j + 1];
        }
        const uint32_t g = (uint32_t)(e >> 32);
        if (g != cur) {
            if (first && off[cur] < j0)
                stp(&head[l], acc);
            else
                stp(&buckets[cur], acc);
            xyzz_set_inf(acc);
            cur = g;
            first = false;
        }
        if (aff_is_inf(pt)) continue;
        if ((uint32_t)e >> 31) O::neg_canon(pt.y, pt.y);  // resident bases are canonical
        xyzz_add_aff<O>(acc, pt.x, pt.y);
    }
    if (first && off[cur] < j0)
        stp(&head[l], acc);
    else if (off[cur + 1] > j1)
        stp(&tail[l], acc);
    else
        stp(&buckets[cur], acc);
}

// First and last accumulation lane whose chunk holds entries of bucket g
// (entries [s, e) of the MSM described by m).
__device__ __forceinline__ void bucket_lanes(const MsmDesc& m, const uint32_t* __restrict__ off, uint32_t s,
                                             uint32_t e, uint32_t& l0, uint32_t& l1) {
    const uint32_t e0 = off[m.bkt_base];
    l0 = m.lane_base + (s - e0) / m.chunk;
    l1 = m.lane_base + (e - 1 - e0) / m.chunk;
}

// One lane per bucket: empty buckets become O; a bucket cut across chunks is
// tail[first lane] + head[every later lane it reaches]. A bucket cut across
// more than HEAVY_SPAN chunks (few, very full buckets: a top window with few
// bits, or repeated scalars) is queued for k_bucket_fixup_heavy instead of
// being summed serially by one lane. The first lane also copies the batch's
// entry count (the scan's total) next to the error flag for the host.
constexpr uint32_t HEAVY_SPAN = 8;

template <class O>
__global__ void __launch_bounds__(64) k_bucket_fixup(const MsmDesc* __restrict__ desc, int count, int nbk,
                                                     const uint32_t* __restrict__ off,
                                                     const typename Stor<O>::X* __restrict__ head,
                                                     const typename Stor<O>::X* __restrict__ tail, typename Stor<O>::X* __restrict__ buckets,
                                                     uint32_t* __restrict__ heavy, uint32_t* __restrict__ heavy_count,
                                                     uint32_t* __restrict__ entries) {
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) / Stor<O>::LPI;
    if (blockIdx.x == 0 && threadIdx.x == 0) *entries = off[nbk];
    if (g >= nbk) return;
    const uint32_t s = off[g], e = off[g + 1];
    Xyzz<O> acc;
    if (s == e) {
        xyzz_set_inf(acc);
        stp(&buckets[g], acc);
        return;
    }
    uint32_t l0, l1;
    bucket_lanes(desc[GM_FIND(desc, count, bkt_base, (uint32_t)g)], off, s, e, l0, l1);
    if (l0 == l1) return;
    if (l1 - l0 > HEAVY_SPAN) {
        if (threadIdx.x % Stor<O>::LPI == 0) heavy[atomicAdd(heavy_count, 1u)] = (uint32_t)g;
        return;
    }
    ldp(acc, &tail[l0]);
    for (uint32_t l = l0 + 1; l <= l1; ++l) {
        Xyzz<O> v;
        ldp(v, &head[l]);
        xyzz_add<O>(acc, acc, v);
    }
    stp(&buckets[g], acc);
}

// Sums sh[0..T) into sh[0..LPI) by a tree in LDS. The first `items` items
// (LPI threads each) hold values, the rest O, so levels above the first power
// of two >= items are skipped; the test is uniform over the block, and the
// levels are unrolled at compile time. The caller has stored its values and
// synchronised.
template <class O, int T>
__device__ __forceinline__ void lds_tree(Xyzz<O>* sh, int tid, uint32_t items) {
    uint32_t span = 1;
    while (span < items) span <<= 1;
    const int lim = (int)span * Stor<O>::LPI;
#pragma unroll
    for (int s = T / 2; s >= Stor<O>::LPI; s >>= 1) {
        if (s < lim) {
            if (tid < s) {
                Xyzz<O> a = sh[tid], b = sh[tid + s];
                xyzz_add<O>(a, a, b);
                sh[tid] = a;
            }
            __syncthreads();
        }
    }
}

// One workgroup per queued heavy bucket (grid-stride over the queue): strided
// per-thread sums of its chunk partials, then a tree in LDS.
template <class O, int THREADS>
__global__ void __launch_bounds__(THREADS) k_bucket_fixup_heavy(const MsmDesc* __restrict__ desc, int count,
                                                                const uint32_t* __restrict__ off,
                                                                const typename Stor<O>::X* __restrict__ head,
                                                                const typename Stor<O>::X* __restrict__ tail,
                                                                typename Stor<O>::X* __restrict__ buckets,
                                                                const uint32_t* __restrict__ heavy,
                                                                const uint32_t* __restrict__ heavy_count) {
    __shared__ Xyzz<O> sh[THREADS];
    const uint32_t nheavy = *heavy_count;
    const int tid = threadIdx.x;
    for (uint32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const uint32_t g = heavy[h];
        uint32_t l0, l1;
        bucket_lanes(desc[GM_FIND(desc, count, bkt_base, g)], off, off[g], off[g + 1], l0, l1);
        const uint32_t parts = l1 - l0 + 1;  // part 0 = tail[l0], part i = head[l0 + i]
        Xyzz<O> acc;
        xyzz_set_inf(acc);
        for (uint32_t i = tid / Stor<O>::LPI; i < parts; i += THREADS / Stor<O>::LPI) {
            Xyzz<O> v;
            if (i == 0)
                ldp(v, &tail[l0]);
            else
                ldp(v, &head[l0 + i]);
            xyzz_add<O>(acc, acc, v);
        }
        sh[tid] = acc;
        __syncthreads();
        lds_tree<O, THREADS>(sh, tid, min((uint32_t)(THREADS / Stor<O>::LPI), parts));
        if (tid < Stor<O>::LPI) stp(&buckets[g], sh[tid]);
        __syncthreads();
    }
}

// ----------------------------------------------------------- bucket reduce

// One lane per (MSM, window, segment of seg buckets). Bucket b of a window
// holds digit magnitude b+1. Output: sum_{b in segment} (b+1) B_b, computed
// as T + lo * S from running sums S = sum B_b, T = sum (b - lo + 1) B_b.
template <class O>
__global__ void __launch_bounds__(64) k_bucket_reduce(const MsmDesc* __restrict__ desc, int count, uint32_t seg_total,
                                                      const typename Stor<O>::X* __restrict__ buckets,
                                                      typename Stor<O>::X* __restrict__ segs) {
    const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) / Stor<O>::LPI;
    if (g >= seg_total) return;
    const MsmDesc& m = desc[GM_FIND(desc, count, seg_base, g)];
    const uint32_t local = g - m.seg_base;
    const uint32_t w = local / m.nseg, t = local - w * m.nseg;
    const typename Stor<O>::X* bw = buckets + m.bkt_base + (size_t)w * m.B + (size_t)t * m.seg;
    const uint32_t lo = t * m.seg;
    const int len = (int)min(m.seg, m.B - lo);  // the window's last segment may be shorter
    Xyzz<O> S, T;
    xyzz_set_inf(S);
    xyzz_set_inf(T);
    for (int b = len - 1; b >= 0; --b) {
        Xyzz<O> bk;
        ldp(bk, &bw[b]);
        xyzz_add<O>(S, S, bk);
        xyzz_add<O>(T, T, S);
    }
    if (lo != 0 && !xyzz_is_inf(S)) {
        Xyzz<O> M;
        xyzz_set_inf(M);
        const int top = 31 - __builtin_clz(lo);
        for (int b = top; b >= 0; --b) {
            xyzz_dbl<O>(M, M);
            if ((lo >> b) & 1) xyzz_add<O>(M, M, S);
        }
        xyzz_add<O>(T, T, M);
    }
    stp(&segs[g], T);
}

// Window sums in two short stages, so the serial depth stays small and the
// whole GPU takes part (one workgroup per window would leave most CUs idle):
//   k_window_partial  one 64-thread workgroup per tile of WSUM_TILE segments
//                     of one window: up to 4 strided adds per thread, then a
//                     tree in LDS over the items that hold segments;
//   k_window_sum      one workgroup per (MSM, window): tree over its tiles.
// Windows of MSMs flagged DESC_WSUM_DIRECT (few segments) skip both and are
// summed by k_window_direct, one lane per window.
template <class O>
__global__ void __launch_bounds__(64 * Stor<O>::LPI) k_window_partial(const MsmDesc* __restrict__ desc, int count,
                                                       const typename Stor<O>::X* __restrict__ segs,
                                                       typename Stor<O>::X* __restrict__ tiles) {
    constexpr int T = 64 * Stor<O>::LPI;
    __shared__ Xyzz<O> sh[T];
    const uint32_t gt = blockIdx.x;
    const int tid = threadIdx.x;
    const MsmDesc& m = desc[GM_FIND(desc, count, tile_base, gt)];
    const uint32_t ntile = (m.nseg + WSUM_TILE - 1) / WSUM_TILE;
    const uint32_t local = gt - m.tile_base;
    const uint32_t w = local / ntile, k = local - w * ntile;
    const typename Stor<O>::X* sw = segs + m.seg_base + (size_t)w * m.nseg;
    const uint32_t t1 = min((k + 1) * WSUM_TILE, m.nseg);
    Xyzz<O> acc;
    xyzz_set_inf(acc);
    for (uint32_t t = k * WSUM_TILE + tid / Stor<O>::LPI; t < t1; t += T / Stor<O>::LPI) {
        Xyzz<O> v;
        ldp(v, &sw[t]);
        xyzz_add<O>(acc, acc, v);
    }
    sh[tid] = acc;
    __syncthreads();
    lds_tree<O, T>(sh, tid, min((uint32_t)(T / Stor<O>::LPI), t1 - k * WSUM_TILE));
    if (tid < Stor<O>::LPI) stp(&tiles[gt], sh[tid]);
}

template <class O, int THREADS>
__global__ void __launch_bounds__(THREADS) k_window_sum(const MsmDesc* __restrict__ desc, int count,
                                                        const typename Stor<O>::X* __restrict__ tiles,
                                                        typename Stor<O>::X* __restrict__ wsum) {
    __shared__ Xyzz<O> sh[THREADS];
    const uint32_t gw = blockIdx.x;
    const int tid = threadIdx.x;
    const MsmDesc& m = desc[GM_FIND(desc, count, win_base, gw)];
    if (m.flags & DESC_WSUM_DIRECT) return;  // k_window_direct sums this window
    const uint32_t w = gw - m.win_base;
    const uint32_t ntile = (m.nseg + WSUM_TILE - 1) / WSUM_TILE;
    const typename Stor<O>::X* tw = tiles + m.tile_base + (size_t)w * ntile;
    Xyzz<O> acc;
    xyzz_set_inf(acc);
    for (uint32_t t = tid / Stor<O>::LPI; t < ntile; t += THREADS / Stor<O>::LPI) {
        Xyzz<O> v;
        ldp(v, &tw[t]);
        xyzz_add<O>(acc, acc, v);
    }
    sh[tid] = acc;
    __syncthreads();
    lds_tree<O, THREADS>(sh, tid, min((uint32_t)(THREADS / Stor<O>::LPI), ntile));
    if (tid < Stor<O>::LPI) stp(&wsum[gw], sh[tid]);
}

// Windows with few segments (DESC_WSUM_DIRECT): one lane per window adds its
// segment results in order.
template <class O>
__global__ void __launch_bounds__(64) k_window_direct(const MsmDesc* __restrict__ desc, int count, uint32_t win_total,
                                                      const typename Stor<O>::X* __restrict__ segs,
                                                      typename Stor<O>::X* __restrict__ wsum) {
    const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / Stor<O>::LPI;
    if (gw >= win_total) return;
    const MsmDesc& m = desc[GM_FIND(desc, count, win_base, gw)];
    if (!(m.flags & DESC_WSUM_DIRECT)) return;
    const typename Stor<O>::X* sw = segs + m.seg_base + (size_t)(gw - m.win_base) * m.nseg;
    Xyzz<O> acc;
    ldp(acc, &sw[0]);
    for (uint32_t t = 1; t < m.nseg; ++t) {
        Xyzz<O> v;
        ldp(v, &sw[t]);
        xyzz_add<O>(acc, acc, v);
    }
    stp(&wsum[gw], acc);
}

// Horner over window sums on the device, one lane per MSM; writes the
// canonical affine result of each MSM.
template <class O>
__global__ void __launch_bounds__(64) k_horner(const MsmDesc* __restrict__ desc, int count,
                                               const Xyzz<O>* __restrict__ wsum, Aff<O>* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    const MsmDesc& m = desc[j];
    const Xyzz<O>* ws = wsum + m.win_base;
    Xyzz<O> r;
    ld16(r, &ws[m.nw - 1]);
    for (int w = (int)m.nw - 2; w >= 0; --w) {
        for (uint32_t k = 0; k < m.c; ++k) xyzz_dbl<O>(r, r);
        Xyzz<O> v;
        ld16(v, &ws[w]);
        xyzz_add<O>(r, r, v);
    }
    Aff<O> a;
    if (!xyzz_to_aff<O>(a.x, a.y, r)) aff_set_inf(a);
    st16(&out[j], a);
}

// device affine -> ABI projective (X, Y, Z) = (x, y, R mod p) or (0, 0, 0),
// the representation reference src/curve.cu:396-432 gives an affine point.
template <class O>
__global__ void __launch_bounds__(64) k_aff_to_abi_proj(uint8_t* __restrict__ out, const Aff<O>* __restrict__ in,
                                                        int count) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= count) return;
    using L = AbiLayout<O>;
    Aff<O> p;
    ld16(p, &in[m]);
    uint64_t* o = reinterpret_cast<uint64_t*>(out + (size_t)m * L::PROJ_BYTES);
    for (int j = 0; j < 3 * L::ELEM_U64; ++j) o[j] = 0;
    if (aff_is_inf(p)) return;
    export_elem(o, p.x);
    export_elem(o + L::ELEM_U64, p.y);
    typename O::F one;
    O::one(one);
    export_elem(o + 2 * L::ELEM_U64, one);
}

}  // namespace gm
