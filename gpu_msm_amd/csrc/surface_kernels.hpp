// Device kernels behind the reference's single-operation device-pointer APIs:
//   * point operations on ABI affine records (reference src/curve.cu:2386-2525
//     launches them; the formulas are the affine ones of curve.cu:72-268);
//   * in-place Montgomery conversion of point batches (curve.cu:2343-2352);
//   * Fp / Fp2 element operations (reference src/fp_kernels.cu and
//     src/fp2_kernels.cu launch them; the formulas are src/fp.cu and
//     src/fp2.cu).
// Records stay in the ABI layout: 7 x u64 little endian per Fp, and for points
// x, y followed by a word whose low byte is the `bool infinity`. Arithmetic
// runs on the 16 x 28-bit device field; every result is written canonical
// (< p), which is what the reference's canonical-input arithmetic returns, so
// results match bit for bit. Inputs are taken as the reference assumes them:
// field values below p.
#pragma once
#include <type_traits>

#include "msm_kernels.hpp"
#include "surface.hpp"

namespace gm {

// ------------------------------------------------ Montgomery domain helpers

__device__ __forceinline__ void to_mont(Fp& r, const Fp& a) {
    Fp r2;
    fp_set_const(r2, R2L);
    fp_mul(r, a, r2);
}
__device__ __forceinline__ void from_mont(Fp& r, const Fp& a) {
    Fp one;
    fp_set_zero(one);
    one.v[0] = 1;
    fp_mul(r, a, one);
}
__device__ __forceinline__ void to_mont(Fp2& r, const Fp2& a) {
    to_mont(r.c0, a.c0);
    to_mont(r.c1, a.c1);
}
__device__ __forceinline__ void from_mont(Fp2& r, const Fp2& a) {
    from_mont(r.c0, a.c0);
    from_mont(r.c1, a.c1);
}

// ------------------------------------------------------ ABI point records

template <class O>
struct Rec {
    using F = typename O::F;
    static constexpr int E = AbiLayout<O>::ELEM_U64;

    __device__ static bool inf(const uint64_t* r) { return (r[2 * E] & 0xff) != 0; }
    __device__ static void load(F& x, F& y, const uint64_t* a) {
        import_elem(x, a, false);
        import_elem(y, a + E, false);
    }
    __device__ static void store(uint64_t* r, const F& x, const F& y) {
        export_elem(r, x);
        export_elem(r + E, y);
        r[2 * E] = 0;
    }
    // reference g1/g2_point_at_infinity (src/curve.cu:365-376): x = y = 0, flag set
    __device__ static void store_inf(uint64_t* r) {
        for (int j = 0; j < 2 * E; ++j) r[j] = 0;
        r[2 * E] = 1;
    }
    __device__ static void copy(uint64_t* r, const uint64_t* a) {
        const uint64_t f = inf(a) ? 1 : 0;
        for (int j = 0; j < 2 * E; ++j) r[j] = a[j];
        r[2 * E] = f;
    }
    __device__ static void store_xyzz(uint64_t* r, const Xyzz<O>& p) {
        F x, y;
        if (xyzz_to_aff<O>(x, y, p))
            store(r, x, y);
        else
            store_inf(r);
    }
};

// One point operation, one lane. r may alias a or b: every input is read
// before the first store.
template <class O>
__global__ void __launch_bounds__(64) k_point_op(int op, uint64_t* r, const uint64_t* a, const uint64_t* b,
                                                 const uint64_t* d_sc, uint64_t sc, int limbs) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    using R = Rec<O>;
    using F = typename O::F;
    constexpr int E = R::E;
    switch (op) {
        case PO_INF:
            R::store_inf(r);
            return;
        case PO_ADD: {  // reference src/curve.cu:173-227
            if (R::inf(a)) {
                R::copy(r, b);
                return;
            }
            if (R::inf(b)) {
                R::copy(r, a);
                return;
            }
            F x1, y1, x2, y2;
            R::load(x1, y1, a);
            R::load(x2, y2, b);
            Xyzz<O> acc;
            xyzz_set_inf(acc);
            xyzz_add_aff<O>(acc, x1, y1);
            xyzz_add_aff<O>(acc, x2, y2);  // equal points double, opposite points give O
            R::store_xyzz(r, acc);
            return;
        }
        case PO_DOUBLE: {  // reference src/curve.cu:85-120: O and y = 0 give O
            if (R::inf(a)) {
                R::store_inf(r);
                return;
            }
            F x, y;
            R::load(x, y, a);
            Xyzz<O> acc;
            xyzz_dbl_aff<O>(acc, x, y);  // y = 0 makes ZZ = 0
            R::store_xyzz(r, acc);
            return;
        }
        case PO_NEG: {  // reference src/curve.cu:72-83: x copied, y -> p - y
            if (R::inf(a)) {
                R::store_inf(r);
                return;
            }
            F y;
            import_elem(y, a + E, false);
            O::neg(y, y);
            uint64_t xw[E];
            for (int j = 0; j < E; ++j) xw[j] = a[j];
            for (int j = 0; j < E; ++j) r[j] = xw[j];
            export_elem(r + E, y);
            r[2 * E] = 0;
            return;
        }
        case PO_SMUL: {  // reference src/curve.cu:231-268: left-to-right double and add
            if (R::inf(a)) {
                R::store_inf(r);
                return;
            }
            F x, y;
            R::load(x, y, a);
            Xyzz<O> acc;
            xyzz_set_inf(acc);
            for (int l = limbs - 1; l >= 0; --l) {
                const uint64_t w = d_sc ? d_sc[l] : sc;
                for (int bit = 63; bit >= 0; --bit) {
                    xyzz_dbl<O>(acc, acc);
                    if ((w >> bit) & 1) xyzz_add_aff<O>(acc, x, y);
                }
            }
            R::store_xyzz(r, acc);
            return;
        }
        case PO_TO_MONT:
        case PO_FROM_MONT: {  // reference src/curve.cu:2305-2331
            if (R::inf(a)) {
                R::store_inf(r);
                return;
            }
            F x, y;
            R::load(x, y, a);
            if (op == PO_TO_MONT) {
                to_mont(x, x);
                to_mont(y, y);
            } else {
                from_mont(x, x);
                from_mont(y, y);
            }
            R::store(r, x, y);
            return;
        }
        default:
            return;
    }
}

// In-place conversion of n ABI records; records at infinity are left as they
// are (reference src/curve.cu:2333-2352).
template <class O>
__global__ void __launch_bounds__(128) k_point_mont_batch(uint64_t* pts, int n, int to) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using R = Rec<O>;
    constexpr int E = R::E;
    uint64_t* rec = pts + (size_t)i * (2 * E + 1);
    if (R::inf(rec)) return;
    typename O::F x, y;
    R::load(x, y, rec);
    if (to) {
        to_mont(x, x);
        to_mont(y, y);
    } else {
        from_mont(x, x);
        from_mont(y, y);
    }
    export_elem(rec, x);
    export_elem(rec + E, y);
}

// ---------------------------------------------------------- field elements

template <class O>
__device__ __forceinline__ int raw_cmp(const uint64_t* a, const uint64_t* b) {
    // Fp: most significant limb first; Fp2: c0 decides, then c1 (reference
    // src/fp.cu:92-98, src/fp2.cu:3-7)
    constexpr int E = AbiLayout<O>::ELEM_U64;
    for (int part = 0; part < E / 7; ++part)
        for (int j = 6; j >= 0; --j) {
            const uint64_t x = a[part * 7 + j], y = b[part * 7 + j];
            if (x != y) return x < y ? -1 : 1;
        }
    return 0;
}

// Element-wise operation over n elements (n = 1 for the single-element API).
template <class O>
__global__ void __launch_bounds__(128) k_field_op(int op, uint64_t* c, const uint64_t* a, const uint64_t* b, int n,
                                                  uint64_t arg, int* iout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using F = typename O::F;
    constexpr int E = AbiLayout<O>::ELEM_U64;
    constexpr bool IS_FP = std::is_same<O, FpOps>::value;
    const uint64_t* ai = a ? a + (size_t)i * E : nullptr;
    const uint64_t* bi = b ? b + (size_t)i * E : nullptr;
    uint64_t* ci = c ? c + (size_t)i * E : nullptr;
    F x, y, t;
    switch (op) {
        case FO_CMP:
            iout[i] = raw_cmp<O>(ai, bi);
            return;
        case FO_IS_ZERO:
        case FO_IS_ONE: {
            uint64_t acc = 0;
            for (int j = 0; j < E; ++j) acc |= ai[j] ^ ((op == FO_IS_ONE && j == 0) ? 1ull : 0ull);
            iout[i] = acc == 0;
            return;
        }
        case FO_COPY: {
            uint64_t w[E];
            for (int j = 0; j < E; ++j) w[j] = ai[j];
            for (int j = 0; j < E; ++j) ci[j] = w[j];
            return;
        }
        case FO_ZERO:
        case FO_ONE:
            for (int j = 0; j < E; ++j) ci[j] = (op == FO_ONE && j == 0) ? 1 : 0;
            return;
        case FO_CMOV: {
            const uint64_t m = 0 - (arg & 1);
            for (int j = 0; j < E; ++j) ci[j] = (ci[j] & ~m) | (ai[j] & m);
            return;
        }
        default:
            break;
    }
    import_elem(x, ai, false);
    if (bi) import_elem(y, bi, false);
    switch (op) {
        case FO_ADD:
            O::add(t, x, y);
            break;
        case FO_SUB:
            O::sub(t, x, y);
            break;
        case FO_NEG:
            O::neg(t, x);
            break;
        case FO_MONT_MUL:
            O::mul(t, x, y);
            break;
        case FO_MUL:
        case FO_SQUARE:  // (x y R^-1) R^2 R^-1 = x y
            O::mul(t, x, op == FO_SQUARE ? x : y);
            to_mont(t, t);
            break;
        case FO_TO_MONT:
            to_mont(t, x);
            break;
        case FO_FROM_MONT:
            from_mont(t, x);
            break;
        case FO_INV:
        case FO_DIV: {  // normal form: into the Montgomery domain and back
            F yi = op == FO_INV ? x : y;
            to_mont(yi, yi);
            O::inv(yi, yi);
            from_mont(yi, yi);
            if (op == FO_INV) {
                t = yi;
            } else {
                O::mul(t, x, yi);
                to_mont(t, t);
            }
            break;
        }
        case FO_CONJ:
            if constexpr (!IS_FP) {
                t.c0 = x.c0;
                fp_neg(t.c1, x.c1);
                break;
            }
            return;
        case FO_MUL_BY_I:  // i (a0 + a1 i) = -a1 + a0 i
            if constexpr (!IS_FP) {
                fp_neg(t.c0, x.c1);
                t.c1 = x.c0;
                break;
            }
            return;
        case FO_SQRT:
        case FO_IS_QR:
        case FO_POW_U64:
            if constexpr (IS_FP) {
                Fp xm, e;
                to_mont(xm, x);
                if (op == FO_POW_U64) {
                    const uint32_t ew[2] = {(uint32_t)(arg >> 32), (uint32_t)arg};
                    fp_pow_words(e, xm, ew, 2);
                    from_mont(t, e);
                    break;
                }
                bool qr = true;
                if (!fp_is_zero(x)) {
                    fp_pow_words(e, xm, EULER_WORDS, 14);
                    Fp one;
                    fp_set_one(one);
                    qr = fp_eq(e, one);
                }
                if (op == FO_IS_QR) {
                    iout[i] = qr;
                    return;
                }
                iout[i] = qr;
                if (qr) {
                    fp_pow_words(e, xm, SQRT_WORDS, 14);
                    from_mont(t, e);
                } else {
                    fp_set_zero(t);
                }
                break;
            }
            return;
        default:
            return;
    }
    export_elem(ci, t);
}

}  // namespace gm
