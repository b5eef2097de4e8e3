// Point wire formats and validation on the device, for G1 (O = FpOps) and
// G2 (O = Fp2Ops), one lane per point:
//   k_fmt_serialize    ABI affine records (Montgomery) -> byte strings in the
//                      layout of ark-serialize's CanonicalSerialize for
//                      short-Weierstrass affine points, the format tfhe-zk-pok
//                      (the reference's interop target, rust-api/src/
//                      conversions.rs) reads and writes: each Fp is its
//                      canonical integer in 56 little-endian bytes, Fp2 is c0
//                      then c1; compressed = x with flags, uncompressed = x then
//                      y with flags; flags in the top two bits of the last
//                      byte: 0x40 point at infinity (coordinates zero), 0x80 y
//                      is the larger of {y, -y} (integers compared for Fp, c1
//                      first and then c0 for Fp2);
//   k_fmt_deserialize  the inverse, with the square root for compressed input
//                      and optional on-curve / subgroup validation;
//   k_fmt_check        on-curve (y^2 = x^3 + b with the curve's actual b: 1 on
//                      G1, 1 + i on G2) and prime-order subgroup (r P = O)
//                      checks of ABI records;
//   k_fmt_canonical    Montgomery <-> canonical standard-form limbs (what
//                      from_sign_and_limbs takes and into_bigint returns),
//                      infinity records kept.
// Status per point: 0 ok, 1 invalid encoding (a field element >= p, or both
// flag bits), 2 not on the curve (no square root for compressed input), 3 not
// in the prime-order subgroup.
#pragma once
#include "surface_kernels.hpp"

namespace gm {

enum FmtStatus : int { FMT_OK = 0, FMT_BAD_ENCODING = 1, FMT_NOT_ON_CURVE = 2, FMT_NOT_IN_SUBGROUP = 3 };
constexpr int FMT_CHECK_ON_CURVE = 1, FMT_CHECK_SUBGROUP = 2;

// r, the prime order of G1 and G2, in 32-bit words, most significant first.
constexpr uint32_t R_ORDER_WORDS[10] = {0x00000511, 0xb70539f2, 0x7995b349, 0x95830fa4, 0xd04c98cc,
                                        0xc4c050bc, 0x7bb9b0e8, 0xd8ca3461, 0x04280014, 0x00040001};
constexpr uint64_t P_U64[7] = {0x311c0026aab0aaabull, 0x56ee4528c573b5ccull, 0x824e6dc3e23acdeeull,
                               0x0f75a64bbac71602ull, 0x0095a4b78a02fe32ull, 0x200fc34965aad640ull,
                               0x3cdee0fb28c5e535ull};

__device__ __forceinline__ bool u64x7_below_p(const uint64_t* w) {
    for (int j = 6; j >= 0; --j)
        if (w[j] != P_U64[j]) return w[j] < P_U64[j];
    return false;
}

// Canonical value above (p - 1) / 2.
__device__ __forceinline__ bool fp_above_half(const Fp& a) {
    for (int i = NL - 1; i >= 0; --i)
        if (a.v[i] != HALFL[i]) return a.v[i] > HALFL[i];
    return false;
}
__device__ __forceinline__ bool fp_raw_eq(const Fp& a, const Fp& b) {
    uint32_t d = 0;
    for (int i = 0; i < NL; ++i) d |= a.v[i] ^ b.v[i];
    return d == 0;
}

// y is the larger of {y, -y}; y canonical and in normal form.
__device__ __forceinline__ bool elem_larger(const Fp& y) { return fp_above_half(y); }
__device__ __forceinline__ bool elem_larger(const Fp2& y) {
    Fp z;
    fp_set_zero(z);
    return fp_raw_eq(y.c1, z) ? fp_above_half(y.c0) : fp_above_half(y.c1);
}

__device__ __forceinline__ void fp2_pow_words(Fp2& r, const Fp2& a, const uint32_t* E, int nw) {
    Fp2 acc;
    fp_set_one(acc.c0);
    fp_set_zero(acc.c1);
    for (int w = 0; w < nw; ++w)
        for (int b = 31; b >= 0; --b) {
            fp2_sqr(acc, acc);
            if ((E[w] >> b) & 1) fp2_mul(acc, acc, a);
        }
    r = acc;
}

// Square roots in the Montgomery domain; false when a is not a square.
__device__ inline bool field_sqrt(Fp& r, const Fp& a) {
    Fp y, y2;
    fp_pow_words(y, a, SQRT_WORDS, 14);
    fp_sqr(y2, y);
    r = y;
    return fp_eq(y2, a);
}
// Adj and Rodriguez-Henriquez (2012), Algorithm 9 (q = p = 3 mod 4).
__device__ inline bool field_sqrt(Fp2& r, const Fp2& a) {
    // (the algorithm's early exit for non-squares, alpha^(q+1) = -1, is
    // replaced by the final check x^2 = a)
    Fp2 a1, alpha, x0, t, x;
    fp2_pow_words(a1, a, P34_WORDS, 14);
    fp2_mul(t, a1, a);
    fp2_mul(alpha, a1, t);  // a^((q-1)/2)
    Fp m1;
    fp_set_one(m1);
    fp_neg(m1, m1);
    x0 = t;  // a1 a
    Fp z;
    fp_set_zero(z);
    if (fp_eq(alpha.c0, m1) && fp_eq(alpha.c1, z)) {
        fp_neg(x.c0, x0.c1);  // i x0
        x.c1 = x0.c0;
    } else {
        Fp2 b = alpha;
        Fp one;
        fp_set_one(one);
        fp_add(b.c0, b.c0, one);
        fp2_pow_words(b, b, EULER_WORDS, 14);
        fp2_mul(x, b, x0);
    }
    Fp2 x2;
    fp2_sqr(x2, x);
    r = x;
    return fp_eq(x2.c0, a.c0) && fp_eq(x2.c1, a.c1);
}

template <class O>
__device__ __forceinline__ void curve_b(typename O::F& b);
template <>
__device__ __forceinline__ void curve_b<FpOps>(Fp& b) {
    fp_set_one(b);
}
template <>
__device__ __forceinline__ void curve_b<Fp2Ops>(Fp2& b) {
    fp_set_one(b.c0);
    fp_set_one(b.c1);
}

template <class O>
__device__ __forceinline__ bool on_curve(const typename O::F& x, const typename O::F& y) {
    typename O::F l, r, b;
    O::sqr(l, y);
    O::sqr(r, x);
    O::mul(r, r, x);
    curve_b<O>(b);
    O::add(r, r, b);
    O::sub(l, l, r);
    return O::is_zero(l);
}

template <class O>
__device__ __forceinline__ bool in_subgroup(const typename O::F& x, const typename O::F& y) {
    Xyzz<O> acc;
    xyzz_set_inf(acc);
    for (int w = 0; w < 10; ++w)
        for (int b = 31; b >= 0; --b) {
            xyzz_dbl<O>(acc, acc);
            if ((R_ORDER_WORDS[w] >> b) & 1) xyzz_add_aff<O>(acc, x, y);
        }
    return xyzz_is_inf(acc);
}

// Element I/O on u64 words.
__device__ __forceinline__ void put_std(uint64_t* w, const Fp& a_mont) {
    Fp t;
    from_mont(t, a_mont);
    export_elem(w, t);
}
__device__ __forceinline__ void put_std(uint64_t* w, const Fp2& a_mont) {
    put_std(w, a_mont.c0);
    put_std(w + 7, a_mont.c1);
}
// Canonical standard-form words -> Montgomery value; false if a component is
// not below p.
__device__ __forceinline__ bool get_std(Fp& r, const uint64_t* w) {
    if (!u64x7_below_p(w)) return false;
    Fp t;
    fp_from_u64x7(t, w);
    to_mont(r, t);
    return true;
}
__device__ __forceinline__ bool get_std(Fp2& r, const uint64_t* w) {
    const bool a = get_std(r.c0, w);
    const bool b = get_std(r.c1, w + 7);
    return a && b;
}
__device__ __forceinline__ void std_of(Fp& r, const Fp& m) {
    from_mont(r, m);
    fp_canon(r, r);
}
__device__ __forceinline__ void std_of(Fp2& r, const Fp2& m) {
    std_of(r.c0, m.c0);
    std_of(r.c1, m.c1);
}

template <class O>
__global__ void __launch_bounds__(64) k_fmt_serialize(const uint64_t* __restrict__ abi, uint64_t* __restrict__ out,
                                                      int n, int compressed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int E = AbiLayout<O>::ELEM_U64;
    const uint64_t* rec = abi + (size_t)i * (2 * E + 1);
    uint64_t* o = out + (size_t)i * (compressed ? E : 2 * E);
    const int words = compressed ? E : 2 * E;
    if ((rec[2 * E] & 0xff) != 0) {
        for (int j = 0; j < words; ++j) o[j] = 0;
        o[words - 1] = 0x40ull << 56;
        return;
    }
    typename O::F x, y, ys;
    import_elem(x, rec, false);
    import_elem(y, rec + E, false);
    put_std(o, x);
    if (!compressed) put_std(o + E, y);
    std_of(ys, y);
    if (elem_larger(ys)) o[words - 1] |= 0x80ull << 56;
}

template <class O>
__global__ void __launch_bounds__(64) k_fmt_deserialize(const uint64_t* __restrict__ in, uint64_t* __restrict__ abi,
                                                        int n, int compressed, int checks,
                                                        int32_t* __restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using F = typename O::F;
    constexpr int E = AbiLayout<O>::ELEM_U64;
    const int words = compressed ? E : 2 * E;
    uint64_t w[2 * E];
    const uint64_t* src = in + (size_t)i * words;
    for (int j = 0; j < words; ++j) w[j] = src[j];
    const uint32_t fl = (uint32_t)(w[words - 1] >> 56) & 0xC0u;
    w[words - 1] &= ~(0xC0ull << 56);
    uint64_t* rec = abi + (size_t)i * (2 * E + 1);
    auto put_inf = [&]() {
        for (int j = 0; j < 2 * E; ++j) rec[j] = 0;
        rec[2 * E] = 1;
    };
    F x, y;
    bool ok = fl != 0xC0u && get_std(x, w);
    if (ok && !compressed) ok = get_std(y, w + E);
    if (!ok) {
        put_inf();
        status[i] = FMT_BAD_ENCODING;
        return;
    }
    if (fl & 0x40u) {  // infinity (coordinates are not inspected)
        put_inf();
        status[i] = FMT_OK;
        return;
    }
    if (compressed) {
        F rhs, b, ys, yn;
        O::sqr(rhs, x);
        O::mul(rhs, rhs, x);
        curve_b<O>(b);
        O::add(rhs, rhs, b);
        if (!field_sqrt(y, rhs)) {
            put_inf();
            status[i] = FMT_NOT_ON_CURVE;
            return;
        }
        O::neg(yn, y);
        std_of(ys, y);
        // the larger root when the flag says so, the smaller otherwise
        if (elem_larger(ys) != ((fl & 0x80u) != 0)) y = yn;
    } else if ((checks & FMT_CHECK_ON_CURVE) && !on_curve<O>(x, y)) {
        put_inf();
        status[i] = FMT_NOT_ON_CURVE;
        return;
    }
    if ((checks & FMT_CHECK_SUBGROUP) && !in_subgroup<O>(x, y)) {
        put_inf();
        status[i] = FMT_NOT_IN_SUBGROUP;
        return;
    }
    export_elem(rec, x);
    export_elem(rec + E, y);
    rec[2 * E] = 0;
    status[i] = FMT_OK;
}

template <class O>
__global__ void __launch_bounds__(64) k_fmt_check(const uint64_t* __restrict__ abi, int n, int checks,
                                                  int32_t* __restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int E = AbiLayout<O>::ELEM_U64;
    const uint64_t* rec = abi + (size_t)i * (2 * E + 1);
    if ((rec[2 * E] & 0xff) != 0) {
        status[i] = FMT_OK;
        return;
    }
    typename O::F x, y;
    import_elem(x, rec, false);
    import_elem(y, rec + E, false);
    int st = FMT_OK;
    if ((checks & FMT_CHECK_ON_CURVE) && !on_curve<O>(x, y))
        st = FMT_NOT_ON_CURVE;
    else if ((checks & FMT_CHECK_SUBGROUP) && !in_subgroup<O>(x, y))
        st = FMT_NOT_IN_SUBGROUP;
    status[i] = st;
}

template <class O>
__global__ void __launch_bounds__(128) k_fmt_canonical(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                       int n, int to_std) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int E = AbiLayout<O>::ELEM_U64;
    const uint64_t* a = in + (size_t)i * (2 * E + 1);
    uint64_t* r = out + (size_t)i * (2 * E + 1);
    if ((a[2 * E] & 0xff) != 0) {
        for (int j = 0; j < 2 * E; ++j) r[j] = 0;
        r[2 * E] = 1;
        return;
    }
    typename O::F x, y;
    import_elem(x, a, false);
    import_elem(y, a + E, false);
    if (to_std) {
        std_of(x, x);
        std_of(y, y);
    } else {
        to_mont(x, x);
        to_mont(y, y);
    }
    export_elem(r, x);
    export_elem(r + E, y);
    r[2 * E] = 0;
}

}  // namespace gm
