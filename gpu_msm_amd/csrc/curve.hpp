// Device group law for BLS12-446 G1 (over Fp) and G2 (over Fp2), a = 0.
//
// Bucket accumulators use extended Jacobian "XYZZ" coordinates
// (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2), whose mixed addition costs 8M + 2S and
// needs no curve constant. Each formula's final Y3 = A B - C D is one lazily
// reduced difference of products (O::mul_sub_raw), and differences that only
// feed products stay raw (O::sub_raw: no carry normalisation; for a field
// policy without a raw form these are the normalised operations). The point
// at infinity is ZZ = ZZZ = 0. Results are
// compared only after affine normalisation, where every representation of a
// point gives the same canonical (x, y) as the reference's standard-projective
// formulas (reference src/curve.cu:444-778).
//
// Affine inputs use an in-band infinity marker: limb 15 of x set to
// 0xFFFFFFFF, a value no field element in [0, 2p) can take (for G2 in both
// components of x, so either lane of a pair sees it).
#pragma once
#include "field.hpp"

namespace gm {

constexpr uint32_t INF_MARK = 0xFFFFFFFFu;

template <class O>
struct Aff {
    typename O::F x, y;
};

template <class O>
struct Xyzz {
    typename O::F X, Y, ZZ, ZZZ;
};

template <class O>
__device__ __forceinline__ bool aff_is_inf(const Aff<O>& p);
template <>
__device__ __forceinline__ bool aff_is_inf<FpOps>(const Aff<FpOps>& p) {
    return p.x.v[NL - 1] == INF_MARK;
}
template <>
__device__ __forceinline__ bool aff_is_inf<FpAccOps>(const Aff<FpAccOps>& p) {
    return p.x.v[NL - 1] == INF_MARK;
}
template <>
__device__ __forceinline__ bool aff_is_inf<Fp2Ops>(const Aff<Fp2Ops>& p) {
    return p.x.c0.v[NL - 1] == INF_MARK;
}
// Each lane of a pair sees its own component; the marker is in both.
template <>
__device__ __forceinline__ bool aff_is_inf<Fp2LaneOps>(const Aff<Fp2LaneOps>& p) {
    return p.x.v[NL - 1] == INF_MARK;
}
template <>
__device__ __forceinline__ bool aff_is_inf<Fp2LaneAccOps>(const Aff<Fp2LaneAccOps>& p) {
    return p.x.v[NL - 1] == INF_MARK;
}

template <class O>
__device__ __forceinline__ void xyzz_set_inf(Xyzz<O>& p) {
    O::zero(p.X);
    O::zero(p.Y);
    O::zero(p.ZZ);
    O::zero(p.ZZZ);
}

template <class O>
__device__ __forceinline__ bool xyzz_is_inf(const Xyzz<O>& p) {
    return O::is_zero(p.ZZ);
}

// X3 = a^2 - b - 2c (a raw) and X3 = a^2 - 2c: one fused reduction where the
// field policy has it (O::RAW), else a square and normalised subtractions
// through the scratch t.
template <class O>
__device__ __forceinline__ void x3_sqr_sub2(typename O::F& r, const typename O::F& a, const typename O::F& b,
                                            const typename O::F& c, typename O::F& t) {
    if constexpr (O::RAW) {
        O::sqr_sub2(r, a, b, c);
    } else {
        O::sqr(t, a);
        O::sub(t, t, b);
        O::sub(t, t, c);
        O::sub(r, t, c);
    }
}
template <class O>
__device__ __forceinline__ void x3_sqr_sub2z(typename O::F& r, const typename O::F& a, const typename O::F& c,
                                             typename O::F& t) {
    if constexpr (O::RAW) {
        O::sqr_sub2z(r, a, c);
    } else {
        O::sqr(t, a);
        O::sub(t, t, c);
        O::sub(r, t, c);
    }
}

// 2 * (x, y) for affine input, result in XYZZ (mdbl-2008-s-1).
template <class O>
__device__ __forceinline__ void xyzz_dbl_aff(Xyzz<O>& r, const typename O::F& x, const typename O::F& y) {
    using F = typename O::F;
    F U, V, W, S, M, t;
    O::dbl(U, y);        // U = 2y
    O::sqr(V, U);        // V = U^2
    O::mul(W, U, V);     // W = U V
    O::mul(S, x, V);     // S = x V
    O::sqr(t, x);        // x^2
    O::dbl(M, t);
    O::add(M, M, t);     // M = 3 x^2
    x3_sqr_sub2z<O>(r.X, M, S, t);  // X3 = M^2 - 2S
    O::sub_raw(t, S, r.X);
    O::mul_sub_raw(r.Y, M, t, W, y);  // Y3 = M(S - X3) - W y
    r.ZZ = V;
    r.ZZZ = W;
}

// r = 2p in XYZZ (dbl-2008-s-1).
template <class O>
__device__ __forceinline__ void xyzz_dbl(Xyzz<O>& r, const Xyzz<O>& p) {
    using F = typename O::F;
    if (xyzz_is_inf(p)) {
        r = p;
        return;
    }
    F U, V, W, S, M, t;
    O::dbl(U, p.Y);
    O::sqr(V, U);
    O::mul(W, U, V);
    O::mul(S, p.X, V);
    O::sqr(t, p.X);
    O::dbl(M, t);
    O::add(M, M, t);
    O::mul(r.ZZ, V, p.ZZ);
    O::mul(r.ZZZ, W, p.ZZZ);
    x3_sqr_sub2z<O>(r.X, M, S, t);  // X3 = M^2 - 2S
    O::sub_raw(t, S, r.X);
    O::mul_sub_raw(r.Y, M, t, W, p.Y);  // p.Y is read before r.Y is written (r may alias p)
}

// acc += (x2, y2) affine (madd-2008-s). Handles acc = O, equal points
// (doubling) and opposite points (infinity). The affine input is not O.
template <class O>
__device__ __forceinline__ void xyzz_add_aff(Xyzz<O>& acc, const typename O::F& x2, const typename O::F& y2) {
    using F = typename O::F;
    if (xyzz_is_inf(acc)) {
        acc.X = x2;
        acc.Y = y2;
        O::one(acc.ZZ);
        O::one(acc.ZZZ);
        return;
    }
    F P, R, PP, PPP, Q, t;
    O::mul(P, x2, acc.ZZ);
    O::sub_raw(P, P, acc.X); // P = U2 - X1 (raw: a product operand only)
    O::mul(R, y2, acc.ZZZ);
    O::sub_raw(R, R, acc.Y); // R = S2 - Y1 (raw)
    O::sqr_raw(PP, P);
    if (O::is_zero(PP)) {    // P = 0 (mod p): the same x
        O::sqr_raw(t, R);
        if (O::is_zero(t)) {
            xyzz_dbl_aff<O>(acc, x2, y2);
        } else {
            xyzz_set_inf(acc);
        }
        return;
    }
    O::mul_raw(PPP, PP, P);
    O::mul(Q, acc.X, PP);
    O::mul(acc.ZZ, acc.ZZ, PP);
    O::mul(acc.ZZZ, acc.ZZZ, PPP);
    x3_sqr_sub2<O>(acc.X, R, PPP, Q, t);  // X3 = R^2 - PPP - 2Q
    O::sub_raw(t, Q, acc.X);
    O::mul_sub_raw(acc.Y, R, t, acc.Y, PPP);  // Y3 = R(Q - X3) - Y1 PPP, one reduction
}

// r = p + q, both XYZZ (add-2008-s). r may alias p or q.
template <class O>
__device__ __forceinline__ void xyzz_add(Xyzz<O>& r, const Xyzz<O>& p, const Xyzz<O>& q) {
    using F = typename O::F;
    if (xyzz_is_inf(p)) {
        r = q;
        return;
    }
    if (xyzz_is_inf(q)) {
        r = p;
        return;
    }
    F U1, S1, P, R, PP, PPP, Q, t;
    O::mul(U1, p.X, q.ZZ);
    O::mul(P, q.X, p.ZZ);
    O::sub_raw(P, P, U1);    // P = U2 - U1 (raw: a product operand only)
    O::mul(S1, p.Y, q.ZZZ);
    O::mul(R, q.Y, p.ZZZ);
    O::sub_raw(R, R, S1);    // R = S2 - S1 (raw)
    O::sqr_raw(PP, P);
    if (O::is_zero(PP)) {    // P = 0 (mod p): the same x
        O::sqr_raw(t, R);
        if (O::is_zero(t)) {
            xyzz_dbl<O>(r, p);
        } else {
            xyzz_set_inf(r);
        }
        return;
    }
    O::mul_raw(PPP, PP, P);
    O::mul(Q, U1, PP);
    O::mul(t, p.ZZ, q.ZZ);
    O::mul(r.ZZ, t, PP);
    O::mul(t, p.ZZZ, q.ZZZ);
    O::mul(r.ZZZ, t, PPP);
    x3_sqr_sub2<O>(r.X, R, PPP, Q, t);  // X3 = R^2 - PPP - 2Q
    O::sub_raw(t, Q, r.X);
    O::mul_sub_raw(r.Y, R, t, S1, PPP);
}

// r = -p
template <class O>
__device__ __forceinline__ void xyzz_neg(Xyzz<O>& r, const Xyzz<O>& p) {
    r = p;
    O::neg(r.Y, p.Y);
}

// Affine canonical (x, y) from XYZZ; returns false for infinity.
template <class O>
__device__ __forceinline__ bool xyzz_to_aff(typename O::F& x, typename O::F& y, const Xyzz<O>& p) {
    using F = typename O::F;
    if (xyzz_is_inf(p)) return false;
    F t, iv;
    O::mul(t, p.ZZ, p.ZZZ);
    O::inv(iv, t);                // 1 / (ZZ ZZZ)
    O::mul(t, iv, p.ZZZ);         // 1 / ZZ
    O::mul(x, p.X, t);
    O::mul(t, iv, p.ZZ);          // 1 / ZZZ
    O::mul(y, p.Y, t);
    O::canon(x, x);
    O::canon(y, y);
    return true;
}

}  // namespace gm
