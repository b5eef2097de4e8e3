// Host-callable half of the reference's C++ field and curve surface:
//   include/fp.h:34-148    (Fp element functions on references)
//   include/fp2.h:20-89    (Fp2 element functions on references)
//   include/curve.h:50-134 (curve constants, on-curve checks, affine and
//                           standard-projective point functions, generators)
// The reference compiles these __host__ __device__; a caller of this library
// runs them on the host (device code of another module cannot call into a
// shared library), built on the x86-64 field code of host_field.hpp.
//
// Semantics follow the reference function by function: values are 7 x 64-bit
// little-endian limbs; fp_mul / fp_inv / fp_div / fp_pow / fp_sqrt work on
// normal-form values, fp_mont_* and the point functions on Montgomery images
// (reference src/curve.cu:20-65 binds point arithmetic to fp_mont_mul /
// fp_mont_inv). Inputs below p give canonical results, equal to the
// reference's bit for bit. Two deliberate departures, both corrections:
// curve_b_g1 / curve_b_g2 are the curves' real coefficients 1 and 1 + i (the
// reference's 4 and 4i, src/curve.cu:280-289, reject its own generators), and
// is_on_curve_g1 / _g2 test against them.
#include <string.h>

#include "../../include/curve.h"
#include "../../include/fp.h"
#include "../../include/fp2.h"
#include "../../include/gpu_msm_ext.h"
#include "host_field.hpp"

namespace {

static_assert(sizeof(gmh::Fp) == sizeof(Fp), "limb layout");
static_assert(sizeof(gmh::Fp2) == sizeof(Fp2), "limb layout");

// Values cross between the ABI structs and gmh's by copies (no aliasing
// casts): every function here may be called with its output aliasing an
// input, as the reference allows (tests/test_msm.cu:137-138 converts G.x in
// place).
inline gmh::Fp h(const Fp& a) {
    gmh::Fp r;
    memcpy(&r, &a, sizeof r);
    return r;
}
inline gmh::Fp2 h(const Fp2& a) {
    gmh::Fp2 r;
    memcpy(&r, &a, sizeof r);
    return r;
}
inline void put(Fp& r, const gmh::Fp& v) { memcpy(&r, &v, sizeof r); }
inline void put(Fp2& r, const gmh::Fp2& v) { memcpy(&r, &v, sizeof r); }

// Montgomery image of a small constant k (k * R mod p).
gmh::Fp mont_small(uint64_t k) {
    gmh::Fp v = {{k, 0, 0, 0, 0, 0, 0}}, r;
    gmh::to_mont(r, v);
    return r;
}

// Square-and-multiply over the exponent's bits, most significant first, in
// Montgomery form; base and result in normal form (reference
// src/fp.cu:448-504). A zero exponent gives 1.
void pow_normal(gmh::Fp& r, const gmh::Fp& a, const uint64_t* e, int limbs) {
    gmh::Fp b, acc = gmh::ONE;
    gmh::to_mont(b, a);
    for (int i = limbs - 1; i >= 0; --i)
        for (int k = 63; k >= 0; --k) {
            gmh::sqr(acc, acc);
            if ((e[i] >> k) & 1) gmh::mul(acc, acc, b);
        }
    gmh::from_mont(r, acc);
}

// (p - 1) / 2 and (p + 1) / 4 (p = 3 mod 4).
gmh::Fp p_shifted(int add, int shift) {
    uint64_t t[8] = {0};
    unsigned __int128 c = (unsigned __int128)gmh::P.l[0] + (uint64_t)(int64_t)add;
    // add is -1 or +1: sign-extend across the words
    const uint64_t ext = add < 0 ? ~0ull : 0ull;
    t[0] = (uint64_t)c;
    c >>= 64;
    for (int i = 1; i < 7; ++i) {
        c += (unsigned __int128)gmh::P.l[i] + ext;
        t[i] = (uint64_t)c;
        c >>= 64;
    }
    gmh::Fp r;
    for (int i = 0; i < 7; ++i) r.l[i] = (t[i] >> shift) | (i < 6 ? t[i + 1] << (64 - shift) : 0);
    return r;
}

bool is_one_normal(const gmh::Fp& a) {
    if (a.l[0] != 1) return false;
    for (int i = 1; i < 7; ++i)
        if (a.l[i]) return false;
    return true;
}

// Montgomery reduction of a 14-word value with the reference's final step:
// one subtraction of p when the 449-bit quotient is >= p (reference
// src/fp.cu:306-366); inputs below p * 2^448 give canonical results.
void redc14(gmh::Fp& r, const uint64_t* a) {
    uint64_t t[15];
    memcpy(t, a, 14 * 8);
    t[14] = 0;
    for (int i = 0; i < 7; ++i) {
        const uint64_t m = t[i] * gmh::PINV;
        unsigned __int128 c = 0;
        for (int j = 0; j < 7; ++j) {
            c += (unsigned __int128)m * gmh::P.l[j] + t[i + j];
            t[i + j] = (uint64_t)c;
            c >>= 64;
        }
        for (int k = i + 7; c && k < 15; ++k) {
            c += t[k];
            t[k] = (uint64_t)c;
            c >>= 64;
        }
    }
    memcpy(r.l, t + 7, 56);
    if (t[14] || gmh::geq_p(r.l)) gmh::sub_p(r.l);
}

// ------------------------------------------------- affine points (Montgomery)
// The reference's affine group law (src/curve.cu:72-268): O + P copies P as
// stored, y = 0 doubles to O, O is (0, 0, infinity).

template <class F, class PT>
void aff_inf(PT& r) {
    memset(&r.x, 0, sizeof r.x);
    memset(&r.y, 0, sizeof r.y);
    r.infinity = true;
}

template <class F, class PT>
F X(const PT& p) {
    F r;
    memcpy(&r, &p.x, sizeof r);
    return r;
}
template <class F, class PT>
F Y(const PT& p) {
    F r;
    memcpy(&r, &p.y, sizeof r);
    return r;
}

template <class F, class PT>
void aff_set(PT& r, const F& x, const F& y) {
    memcpy(&r.x, &x, sizeof x);
    memcpy(&r.y, &y, sizeof y);
    r.infinity = false;
}

template <class F, class PT>
void aff_double(PT& r, const PT& p) {
    if (p.infinity || gmh::is_zero(Y<F>(p))) {
        aff_inf<F>(r);
        return;
    }
    F t, n, d, l, x3, y3;
    gmh::sqr(t, X<F>(p));
    gmh::add(n, t, t);
    gmh::add(n, n, t);       // 3 x^2
    gmh::add(d, Y<F>(p), Y<F>(p));
    gmh::inv(d, d);
    gmh::mul(l, n, d);       // lambda
    gmh::sqr(t, l);
    gmh::sub(t, t, X<F>(p));
    gmh::sub(x3, t, X<F>(p));
    gmh::sub(t, X<F>(p), x3);
    gmh::mul(t, l, t);
    gmh::sub(y3, t, Y<F>(p));
    aff_set<F>(r, x3, y3);
}

template <class F, class PT>
void aff_add(PT& r, const PT& a, const PT& b) {
    if (a.infinity) {
        PT t = b;
        r = t;
        return;
    }
    if (b.infinity) {
        PT t = a;
        r = t;
        return;
    }
    if (gmh::eq(X<F>(a), X<F>(b))) {
        F ny;
        gmh::neg(ny, Y<F>(b));
        if (gmh::eq(Y<F>(a), ny)) {
            aff_inf<F>(r);
            return;
        }
        if (gmh::eq(Y<F>(a), Y<F>(b))) {
            PT t = a;
            aff_double<F>(r, t);
            return;
        }
    }
    F dx, dy, l, t, x3, y3;
    gmh::sub(dx, X<F>(b), X<F>(a));
    gmh::sub(dy, Y<F>(b), Y<F>(a));
    gmh::inv(dx, dx);
    gmh::mul(l, dx, dy);
    gmh::sqr(t, l);
    gmh::sub(t, t, X<F>(a));
    gmh::sub(x3, t, X<F>(b));
    gmh::sub(t, X<F>(a), x3);
    gmh::mul(t, l, t);
    gmh::sub(y3, t, Y<F>(a));
    aff_set<F>(r, x3, y3);
}

template <class F, class PT>
void aff_neg(PT& r, const PT& p) {
    if (p.infinity) {
        aff_inf<F>(r);
        return;
    }
    F x = X<F>(p), y;
    gmh::neg(y, Y<F>(p));
    aff_set<F>(r, x, y);
}

// Double-and-add in Jacobian coordinates, one inversion at the end: the same
// canonical affine result as the reference's affine double-and-add
// (src/curve.cu:231-268), whose every step is the group law.
template <class F, class PT>
void aff_smul(PT& r, const PT& p, const uint64_t* k, int limbs) {
    if (p.infinity) {
        aff_inf<F>(r);
        return;
    }
    gmh::Jac<F> acc, base;
    memset(&acc, 0, sizeof acc);
    base.X = X<F>(p);
    base.Y = Y<F>(p);
    memset(&base.Z, 0, sizeof base.Z);
    memcpy(&base.Z, gmh::ONE.l, 56);
    if (gmh::is_zero(base.Y)) {  // (x, 0) has order 2: k P is O or P
        bool odd = limbs > 0 && (k[0] & 1);
        if (odd) {
            PT t = p;
            r = t;
        } else {
            aff_inf<F>(r);
        }
        return;
    }
    for (int i = limbs - 1; i >= 0; --i)
        for (int b = 63; b >= 0; --b) {
            gmh::jdbl(acc, acc);
            if ((k[i] >> b) & 1) gmh::jadd(acc, acc, base);
        }
    F x, y;
    if (!gmh::to_affine(x, y, acc)) {
        aff_inf<F>(r);
        return;
    }
    aff_set<F>(r, x, y);
}

// ------------------------------------- standard projective (X:Y:Z), Montgomery
// The reference's formulas (src/curve.cu:444-738), so the raw (X:Y:Z), not
// only its affine image, matches: add with u = Y2 Z1 - Y1 Z2, v = X2 Z1 - X1 Z2
// (u = v = 0 doubles), X3 = vA, Y3 = u(R - A) - v^3 Y1 Z2, Z3 = v^3 Z1 Z2;
// double with A = 3X^2, B = YZ, C = XYB, D = A^2 - 8C.

template <class F>
struct ProjF {
    F X, Y, Z;
};

template <class F>
F small_c(uint64_t k);
template <>
gmh::Fp small_c<gmh::Fp>(uint64_t k) {
    return mont_small(k);
}
template <>
gmh::Fp2 small_c<gmh::Fp2>(uint64_t k) {
    gmh::Fp2 r;
    r.c0 = mont_small(k);
    memset(&r.c1, 0, sizeof r.c1);
    return r;
}

template <class F>
void proj_double(ProjF<F>& r, const ProjF<F>& p) {
    if (gmh::is_zero(p.Z)) {
        r = p;
        return;
    }
    const F two = small_c<F>(2), three = small_c<F>(3), four = small_c<F>(4), eight = small_c<F>(8);
    F t, A, B, C, D, u, w;
    gmh::mul(t, p.X, p.X);
    gmh::mul(A, three, t);
    gmh::mul(B, p.Y, p.Z);
    gmh::mul(t, p.X, p.Y);
    gmh::mul(C, t, B);
    gmh::mul(t, A, A);
    gmh::mul(u, eight, C);
    gmh::sub(D, t, u);
    ProjF<F> o;
    gmh::mul(t, B, D);
    gmh::mul(o.X, two, t);
    gmh::mul(u, four, C);
    gmh::sub(u, u, D);
    gmh::mul(u, A, u);
    gmh::mul(t, p.Y, p.Y);
    gmh::mul(w, B, B);
    gmh::mul(t, t, w);
    gmh::mul(t, eight, t);
    gmh::sub(o.Y, u, t);
    gmh::mul(t, w, B);
    gmh::mul(o.Z, eight, t);
    r = o;
}

template <class F>
void proj_add(ProjF<F>& r, const ProjF<F>& a, const ProjF<F>& b) {
    if (gmh::is_zero(a.Z)) {
        r = b;
        return;
    }
    if (gmh::is_zero(b.Z)) {
        r = a;
        return;
    }
    F Y1Z2, X1Z2, Z1Z2, u, uu, v, vv, vvv, R, A, t;
    gmh::mul(Y1Z2, a.Y, b.Z);
    gmh::mul(X1Z2, a.X, b.Z);
    gmh::mul(Z1Z2, a.Z, b.Z);
    gmh::mul(t, b.Y, a.Z);
    gmh::sub(u, t, Y1Z2);
    gmh::mul(uu, u, u);
    gmh::mul(t, b.X, a.Z);
    gmh::sub(v, t, X1Z2);
    if (gmh::is_zero(u) && gmh::is_zero(v)) {
        ProjF<F> c = a;
        proj_double(r, c);
        return;
    }
    gmh::mul(vv, v, v);
    gmh::mul(vvv, v, vv);
    gmh::mul(R, vv, X1Z2);
    gmh::mul(t, uu, Z1Z2);
    gmh::sub(t, t, vvv);
    F twoR;
    gmh::add(twoR, R, R);
    gmh::sub(A, t, twoR);
    ProjF<F> o;
    gmh::mul(o.X, v, A);
    gmh::sub(t, R, A);
    gmh::mul(t, u, t);
    F w;
    gmh::mul(w, vvv, Y1Z2);
    gmh::sub(o.Y, t, w);
    gmh::mul(o.Z, vvv, Z1Z2);
    r = o;
}

template <class F, class PR>
ProjF<F> pj(const PR& p) {
    ProjF<F> r;
    memcpy(&r, &p, sizeof r);
    return r;
}
template <class F, class PR>
void pj_put(PR& r, const ProjF<F>& v) {
    memcpy(&r, &v, sizeof r);
}

const G1Point G1_GEN_STANDARD = {
    {{0x3bf9166c8236f4faULL, 0x8bc02b7cbe6a9e8dULL, 0x11c1e56b3e4bc80bULL, 0x6b20d782901a6f62ULL,
      0x2ce8c34265bf3841ULL, 0x11b73d3d76ae9851ULL, 0x326ed6bd777fc6a3ULL}},
    {{0xfe6f792612016b30ULL, 0x22db0ce6034a9db9ULL, 0xb9093f32002756daULL, 0x39d7f424b6660204ULL,
      0xf843c947aa57f571ULL, 0xd6d62d244e413636ULL, 0x1a7caf4a4d3887a6ULL}},
    false};

const G2Point G2_GEN_STANDARD = {
    {{{0x0e529ee4dce9991dULL, 0xd6ebaf149094f1ccULL, 0x043c6bf16312d638ULL, 0x062b61439640e885ULL,
       0x18dad8ed784dd225ULL, 0xa57c0038441f7d15ULL, 0x21f8d4a76f74541aULL}},
     {{0xcaf5185423a7d23aULL, 0x7cef6acb145b6413ULL, 0x2879dd439b019b8bULL, 0x71449cdeca4f0007ULL,
       0xdebaf4a2c5534527ULL, 0xa1b4e791d1b86560ULL, 0x1e0f563c601bb8dcULL}}},
    {{{0x274315837455b919ULL, 0x82039e4221ff3507ULL, 0x00346cebad16a036ULL, 0x0177bfd6654e681eULL,
       0xddff621b5db3f897ULL, 0x0cc61570301497a7ULL, 0x115ea2305a78f646ULL}},
     {{0x392236e9cf2976c2ULL, 0xd8ab17c84b9f03cdULL, 0x8a8e6755f9d82fd1ULL, 0x7532834528cd5a64ULL,
       0x0b0bcc3fb6f2161cULL, 0x76a2ffcb7d47679dULL, 0x25ed2192b203c1feULL}}},
    false};

}  // namespace

// ================================================================ fp.h

int fp_cmp(const Fp& a, const Fp& b) {
    for (int i = FP_LIMBS - 1; i >= 0; --i)
        if (a.limb[i] != b.limb[i]) return a.limb[i] > b.limb[i] ? 1 : -1;
    return 0;
}
bool fp_is_zero(const Fp& a) { return gmh::is_zero(h(a)); }
bool fp_is_one(const Fp& a) { return is_one_normal(h(a)); }
void fp_zero(Fp& a) { memset(&a, 0, sizeof a); }
void fp_one(Fp& a) {
    memset(&a, 0, sizeof a);
    a.limb[0] = 1;
}
void fp_one_montgomery(Fp& a) { put(a, gmh::ONE); }
void fp_copy(Fp& dst, const Fp& src) { memmove(&dst, &src, sizeof dst); }

uint64_t fp_add_raw(Fp& c, const Fp& a, const Fp& b) {
    unsigned __int128 s = 0;
    Fp r;
    for (int i = 0; i < FP_LIMBS; ++i) {
        s += (unsigned __int128)a.limb[i] + b.limb[i];
        r.limb[i] = (uint64_t)s;
        s >>= 64;
    }
    c = r;
    return (uint64_t)s;
}
uint64_t fp_sub_raw(Fp& c, const Fp& a, const Fp& b) {
    uint64_t br = 0;
    Fp r;
    for (int i = 0; i < FP_LIMBS; ++i) {
        unsigned __int128 d = (unsigned __int128)a.limb[i] - b.limb[i] - br;
        r.limb[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) & 1;
    }
    c = r;
    return br;
}
void fp_add(Fp& c, const Fp& a, const Fp& b) {
    gmh::Fp r;
    gmh::add(r, h(a), h(b));
    put(c, r);
}
void fp_sub(Fp& c, const Fp& a, const Fp& b) {
    gmh::Fp r;
    gmh::sub(r, h(a), h(b));
    put(c, r);
}
void fp_mul_raw(uint64_t* c, const Fp& a, const Fp& b) {
    uint64_t t[14] = {0};
    for (int i = 0; i < FP_LIMBS; ++i) {
        unsigned __int128 s = 0;
        for (int j = 0; j < FP_LIMBS; ++j) {
            s += (unsigned __int128)a.limb[i] * b.limb[j] + t[i + j];
            t[i + j] = (uint64_t)s;
            s >>= 64;
        }
        t[i + FP_LIMBS] = (uint64_t)s;
    }
    memcpy(c, t, sizeof t);
}
void fp_mont_reduce(Fp& c, const uint64_t* a) {
    gmh::Fp r;
    redc14(r, a);
    put(c, r);
}
// reference src/fp.cu:300-304: the reduction of a double-width value is the
// Montgomery reduction
void fp_reduce(Fp& c, const uint64_t* a) { fp_mont_reduce(c, a); }
void fp_mont_mul(Fp& c, const Fp& a, const Fp& b) {
    gmh::Fp r;
    gmh::mul(r, h(a), h(b));
    put(c, r);
}
void fp_to_montgomery(Fp& c, const Fp& a) {
    gmh::Fp r;
    gmh::to_mont(r, h(a));
    put(c, r);
}
void fp_from_montgomery(Fp& c, const Fp& a) {
    gmh::Fp r;
    gmh::from_mont(r, h(a));
    put(c, r);
}
void fp_to_montgomery_batch(Fp* dst, const Fp* src, int n) {
    for (int i = 0; i < n; ++i) fp_to_montgomery(dst[i], src[i]);
}
void fp_from_montgomery_batch(Fp* dst, const Fp* src, int n) {
    for (int i = 0; i < n; ++i) fp_from_montgomery(dst[i], src[i]);
}
void fp_mul(Fp& c, const Fp& a, const Fp& b) {
    gmh::Fp am, bm, r;
    gmh::to_mont(am, h(a));
    gmh::to_mont(bm, h(b));
    gmh::mul(r, am, bm);
    gmh::from_mont(r, r);
    put(c, r);
}
void fp_neg(Fp& c, const Fp& a) {
    gmh::Fp r;
    gmh::neg(r, h(a));
    put(c, r);
}
void fp_inv(Fp& c, const Fp& a) {
    gmh::Fp r;
    gmh::to_mont(r, h(a));
    gmh::inv(r, r);  // 0 -> 0
    gmh::from_mont(r, r);
    put(c, r);
}
void fp_mont_inv(Fp& c, const Fp& a) {
    gmh::Fp r;
    gmh::inv(r, h(a));
    put(c, r);
}
void fp_div(Fp& c, const Fp& a, const Fp& b) {
    if (fp_is_zero(b)) {
        fp_zero(c);
        return;
    }
    Fp bi;
    fp_inv(bi, b);
    fp_mul(c, a, bi);
}
void fp_pow(Fp& c, const Fp& a, const uint64_t* e, int e_limbs) {
    const int n = e_limbs > FP_LIMBS ? FP_LIMBS : (e_limbs < 0 ? 0 : e_limbs);
    gmh::Fp r;
    pow_normal(r, h(a), e, n);
    put(c, r);
}
void fp_pow_u64(Fp& c, const Fp& a, uint64_t e) {
    gmh::Fp r;
    pow_normal(r, h(a), &e, 1);
    put(c, r);
}
bool fp_is_quadratic_residue(const Fp& a) {
    if (fp_is_zero(a)) return true;
    const gmh::Fp e = p_shifted(-1, 1);  // (p - 1) / 2: Euler's criterion
    gmh::Fp r;
    pow_normal(r, h(a), e.l, 7);
    return is_one_normal(r);
}
bool fp_sqrt(Fp& c, const Fp& a) {
    if (fp_is_zero(a)) {
        fp_zero(c);
        return true;
    }
    if (!fp_is_quadratic_residue(a)) {
        fp_zero(c);
        return false;
    }
    const gmh::Fp e = p_shifted(1, 2);  // (p + 1) / 4: a^e squares to a for p = 3 mod 4
    gmh::Fp r;
    pow_normal(r, h(a), e.l, 7);
    put(c, r);
    return true;
}
void fp_cmov(Fp& dst, const Fp& src, uint64_t condition) {
    const uint64_t m = 0 - (condition & 1);
    for (int i = 0; i < FP_LIMBS; ++i) dst.limb[i] = (dst.limb[i] & ~m) | (src.limb[i] & m);
}
const Fp& fp_modulus() {
    static const Fp p = {{0x311c0026aab0aaabULL, 0x56ee4528c573b5ccULL, 0x824e6dc3e23acdeeULL, 0x0f75a64bbac71602ULL,
                          0x0095a4b78a02fe32ULL, 0x200fc34965aad640ULL, 0x3cdee0fb28c5e535ULL}};
    return p;
}

// =============================================================== fp2.h

int fp2_cmp(const Fp2& a, const Fp2& b) {
    const int c = fp_cmp(a.c0, b.c0);
    return c ? c : fp_cmp(a.c1, b.c1);
}
bool fp2_is_zero(const Fp2& a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
bool fp2_is_one(const Fp2& a) { return fp_is_one(a.c0) && fp_is_zero(a.c1); }
void fp2_zero(Fp2& a) { memset(&a, 0, sizeof a); }
void fp2_one(Fp2& a) {
    memset(&a, 0, sizeof a);
    a.c0.limb[0] = 1;
}
void fp2_copy(Fp2& dst, const Fp2& src) { memmove(&dst, &src, sizeof dst); }
void fp2_add(Fp2& c, const Fp2& a, const Fp2& b) {
    gmh::Fp2 r;
    gmh::add(r, h(a), h(b));
    put(c, r);
}
void fp2_sub(Fp2& c, const Fp2& a, const Fp2& b) {
    gmh::Fp2 r;
    gmh::sub(r, h(a), h(b));
    put(c, r);
}
void fp2_mont_mul(Fp2& c, const Fp2& a, const Fp2& b) {
    gmh::Fp2 r;
    gmh::mul(r, h(a), h(b));
    put(c, r);
}
void fp2_mul(Fp2& c, const Fp2& a, const Fp2& b) {
    gmh::Fp2 am, bm, r;
    gmh::to_mont(am.c0, h(a.c0));
    gmh::to_mont(am.c1, h(a.c1));
    gmh::to_mont(bm.c0, h(b.c0));
    gmh::to_mont(bm.c1, h(b.c1));
    gmh::mul(r, am, bm);
    gmh::from_mont(r.c0, r.c0);
    gmh::from_mont(r.c1, r.c1);
    put(c, r);
}
void fp2_square(Fp2& c, const Fp2& a) { fp2_mul(c, a, a); }
void fp2_neg(Fp2& c, const Fp2& a) {
    gmh::Fp2 r;
    gmh::neg(r, h(a));
    put(c, r);
}
void fp2_conjugate(Fp2& c, const Fp2& a) {
    gmh::Fp2 r = h(a);
    gmh::neg(r.c1, r.c1);
    put(c, r);
}
void fp2_inv(Fp2& c, const Fp2& a) {
    gmh::Fp2 r;
    gmh::to_mont(r.c0, h(a.c0));
    gmh::to_mont(r.c1, h(a.c1));
    gmh::inv(r, r);  // 0 -> 0
    gmh::from_mont(r.c0, r.c0);
    gmh::from_mont(r.c1, r.c1);
    put(c, r);
}
void fp2_mont_inv(Fp2& c, const Fp2& a) {
    gmh::Fp2 r;
    gmh::inv(r, h(a));
    put(c, r);
}
void fp2_div(Fp2& c, const Fp2& a, const Fp2& b) {
    Fp2 bi;
    fp2_inv(bi, b);
    fp2_mul(c, a, bi);
}
void fp2_cmov(Fp2& dst, const Fp2& src, uint64_t condition) {
    fp_cmov(dst.c0, src.c0, condition);
    fp_cmov(dst.c1, src.c1, condition);
}
void fp2_frobenius(Fp2& c, const Fp2& a) { fp2_conjugate(c, a); }  // i^p = -i for p = 3 mod 4
void fp2_mul_by_i(Fp2& c, const Fp2& a) {
    gmh::Fp2 v = h(a), r;
    gmh::neg(r.c0, v.c1);
    r.c1 = v.c0;
    put(c, r);
}
void fp2_mul_by_non_residue(Fp2& c, const Fp2& a) { fp2_mul_by_i(c, a); }

// ============================================================== curve.h

const Fp& curve_b_g1() {
    static const Fp b = {{1, 0, 0, 0, 0, 0, 0}};
    return b;
}
const Fp2& curve_b_g2() {
    static const Fp2 b = {{{1, 0, 0, 0, 0, 0, 0}}, {{1, 0, 0, 0, 0, 0, 0}}};
    return b;
}
bool is_on_curve_g1(const G1Point& p) {
    if (p.infinity) return true;
    Fp y2, x2, x3, rhs;
    fp_mul(y2, p.y, p.y);
    fp_mul(x2, p.x, p.x);
    fp_mul(x3, x2, p.x);
    fp_add(rhs, x3, curve_b_g1());
    return fp_cmp(y2, rhs) == 0;
}
bool is_on_curve_g2(const G2Point& p) {
    if (p.infinity) return true;
    Fp2 y2, x2, x3, rhs;
    fp2_mul(y2, p.y, p.y);
    fp2_mul(x2, p.x, p.x);
    fp2_mul(x3, x2, p.x);
    fp2_add(rhs, x3, curve_b_g2());
    return fp2_cmp(y2, rhs) == 0;
}
void g1_point_at_infinity(G1Point& p) { aff_inf<gmh::Fp>(p); }
void g2_point_at_infinity(G2Point& p) { aff_inf<gmh::Fp2>(p); }
bool g1_is_infinity(const G1Point& p) { return p.infinity; }
bool g2_is_infinity(const G2Point& p) { return p.infinity; }

template <>
void point_add<G1Point>(G1Point& r, const G1Point& a, const G1Point& b) {
    const G1Point ta = a, tb = b;
    aff_add<gmh::Fp>(r, ta, tb);
}
template <>
void point_add<G2Point>(G2Point& r, const G2Point& a, const G2Point& b) {
    const G2Point ta = a, tb = b;
    aff_add<gmh::Fp2>(r, ta, tb);
}
template <>
void point_double<G1Point>(G1Point& r, const G1Point& p) {
    const G1Point t = p;
    aff_double<gmh::Fp>(r, t);
}
template <>
void point_double<G2Point>(G2Point& r, const G2Point& p) {
    const G2Point t = p;
    aff_double<gmh::Fp2>(r, t);
}
template <>
void point_neg<G1Point>(G1Point& r, const G1Point& p) {
    const G1Point t = p;
    aff_neg<gmh::Fp>(r, t);
}
template <>
void point_neg<G2Point>(G2Point& r, const G2Point& p) {
    const G2Point t = p;
    aff_neg<gmh::Fp2>(r, t);
}
template <>
void point_scalar_mul<G1Point>(G1Point& r, const G1Point& p, const uint64_t* k, int limbs) {
    const G1Point t = p;
    aff_smul<gmh::Fp>(r, t, k, limbs);
}
template <>
void point_scalar_mul<G2Point>(G2Point& r, const G2Point& p, const uint64_t* k, int limbs) {
    const G2Point t = p;
    aff_smul<gmh::Fp2>(r, t, k, limbs);
}

const G1Point& g1_generator() { return G1_GEN_STANDARD; }
const G2Point& g2_generator() { return G2_GEN_STANDARD; }

void g1_projective_point_at_infinity(G1ProjectivePoint& p) { memset(&p, 0, sizeof p); }
void g2_projective_point_at_infinity(G2ProjectivePoint& p) { memset(&p, 0, sizeof p); }

void affine_to_projective(G1ProjectivePoint& proj, const G1Point& a) {
    G1ProjectivePoint r;
    memset(&r, 0, sizeof r);
    if (!a.infinity) {
        r.X = a.x;
        r.Y = a.y;
        put(r.Z, gmh::ONE);
    }
    proj = r;
}
void affine_to_projective(G2ProjectivePoint& proj, const G2Point& a) {
    G2ProjectivePoint r;
    memset(&r, 0, sizeof r);
    if (!a.infinity) {
        r.X = a.x;
        r.Y = a.y;
        put(r.Z.c0, gmh::ONE);
    }
    proj = r;
}
void projective_to_affine_g1(G1Point& a, const G1ProjectivePoint& p) {
    const ProjF<gmh::Fp> v = pj<gmh::Fp>(p);
    if (gmh::is_zero(v.Z)) {
        g1_point_at_infinity(a);
        return;
    }
    gmh::Fp zi, x, y;
    gmh::inv(zi, v.Z);
    gmh::mul(x, v.X, zi);
    gmh::mul(y, v.Y, zi);
    aff_set<gmh::Fp>(a, x, y);
}
void projective_to_affine_g2(G2Point& a, const G2ProjectivePoint& p) {
    const ProjF<gmh::Fp2> v = pj<gmh::Fp2>(p);
    if (gmh::is_zero(v.Z)) {
        g2_point_at_infinity(a);
        return;
    }
    gmh::Fp2 zi, x, y;
    gmh::inv(zi, v.Z);
    gmh::mul(x, v.X, zi);
    gmh::mul(y, v.Y, zi);
    aff_set<gmh::Fp2>(a, x, y);
}
void projective_point_add(G1ProjectivePoint& r, const G1ProjectivePoint& a, const G1ProjectivePoint& b) {
    ProjF<gmh::Fp> o;
    proj_add<gmh::Fp>(o, pj<gmh::Fp>(a), pj<gmh::Fp>(b));
    pj_put<gmh::Fp>(r, o);
}
void projective_point_add(G2ProjectivePoint& r, const G2ProjectivePoint& a, const G2ProjectivePoint& b) {
    ProjF<gmh::Fp2> o;
    proj_add<gmh::Fp2>(o, pj<gmh::Fp2>(a), pj<gmh::Fp2>(b));
    pj_put<gmh::Fp2>(r, o);
}
void projective_point_double(G1ProjectivePoint& r, const G1ProjectivePoint& p) {
    ProjF<gmh::Fp> o;
    proj_double<gmh::Fp>(o, pj<gmh::Fp>(p));
    pj_put<gmh::Fp>(r, o);
}
void projective_point_double(G2ProjectivePoint& r, const G2ProjectivePoint& p) {
    ProjF<gmh::Fp2> o;
    proj_double<gmh::Fp2>(o, pj<gmh::Fp2>(p));
    pj_put<gmh::Fp2>(r, o);
}

// =================================================== extension: closed form

int gm_generator_multiple(int curve, const uint64_t* k, int limbs, void* result_affine) {
    if (!k || !result_affine || limbs < 1 || limbs > 8) return -1;
    if (curve == GM_CURVE_G1) {
        G1Point g = g1_generator(), r;
        memset(&r, 0, sizeof r);
        fp_to_montgomery(g.x, g.x);
        fp_to_montgomery(g.y, g.y);
        aff_smul<gmh::Fp>(r, g, k, limbs);
        memcpy(result_affine, &r, sizeof r);
        return 0;
    }
    if (curve == GM_CURVE_G2) {
        G2Point g = g2_generator(), r;
        memset(&r, 0, sizeof r);
        fp_to_montgomery(g.x.c0, g.x.c0);
        fp_to_montgomery(g.x.c1, g.x.c1);
        fp_to_montgomery(g.y.c0, g.y.c0);
        fp_to_montgomery(g.y.c1, g.y.c1);
        aff_smul<gmh::Fp2>(r, g, k, limbs);
        memcpy(result_affine, &r, sizeof r);
        return 0;
    }
    return -1;
}
