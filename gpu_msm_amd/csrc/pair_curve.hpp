// XYZZ addition and doubling for the split lane groups of the tail kernels:
// FpPairOps (G1: two lanes, each computing Fp products) and Fp2QuadOps (G2:
// two lane pairs, each computing Fp2 products on Fp2LaneOps). Both halves of
// a group hold the same operands and compute two independent products per
// step, so a point operation's serial depth is about half its product count.
// The bucket reduction and window sums run at under one wave per SIMD, where
// a wave's instruction issue (not the SIMD's throughput) bounds their time.
// Representations may differ from the one-group formulas (X3 through
// normalised subtractions instead of the fused square); every value is the
// same point, and results are compared only in affine form.
#pragma once
#include "curve.hpp"

namespace gm {

// F: field operations of one half; half(): which product of a step this
// lane computes; results(r0, r1, mine): both halves' results to every lane.
template <class G>
struct Split;
template <>
struct Split<FpPairOps> {
    using F = FpOps;
    __device__ static __forceinline__ bool half() { return lane_part() != 0; }
    __device__ static __forceinline__ void results(Fp& r0, Fp& r1, const Fp& mine) { pair_results(r0, r1, mine); }
};
template <>
struct Split<Fp2QuadOps> {
    using F = Fp2LaneOps;
    __device__ static __forceinline__ bool half() { return quad_half() != 0; }
    __device__ static __forceinline__ void results(Fp& r0, Fp& r1, const Fp& mine) { quad_results(r0, r1, mine); }
};

// r = 2p (dbl-2008-s-1), four split steps:
//   1: V = U^2 | T = X^2            (U = 2Y)
//   2: W = U V | S = X V            (M = 3T, one fold pass)
//   3: M^2     | ZZ3 = V ZZ         (X3 = M^2 - 2S, one fold pass)
//   4: Y3 = M (S - X3) - W Y | ZZZ3 = W ZZZ   (both as a difference of products)
template <class G>
__device__ __forceinline__ void split_dbl(Xyzz<G>& r, const Xyzz<G>& p) {
    using O = typename Split<G>::F;
    if (xyzz_is_inf(p)) {
        r = p;
        return;
    }
    const bool h = Split<G>::half();
    Fp U, V, T, W, S, M, MM, X3, Z2, Y3, Z3, a, b, c, d, u, zero;
    O::dbl(U, p.Y);
    fp_sel(a, h, p.X, U);
    O::sqr(u, a);
    Split<G>::results(V, T, u);
    fp_sel(a, h, p.X, U);
    O::mul(u, a, V);
    Split<G>::results(W, S, u);
    fp_triple(M, T);
    fp_sel(a, h, V, M);
    fp_sel(b, h, p.ZZ, M);
    O::mul(u, a, b);
    Split<G>::results(MM, Z2, u);
    O::zero(zero);
    fp_sub_b2c(X3, MM, zero, S);
    O::sub_raw(d, S, X3);  // a product operand only
    fp_sel(a, h, W, M);
    fp_sel(b, h, p.ZZZ, d);
    fp_sel(c, h, zero, W);
    fp_sel(d, h, zero, p.Y);
    O::mul_sub_raw(u, a, b, c, d);  // half 0: M (S - X3) - W Y; half 1: W ZZZ - 0
    Split<G>::results(Y3, Z3, u);
    r.X = X3;
    r.Y = Y3;
    r.ZZ = Z2;
    r.ZZZ = Z3;
}

// r = p + q (add-2008-s), seven split steps:
//   1: U1 = X1 ZZ2   | U2 = X2 ZZ1       (P = U2 - U1)
//   2: S1 = Y1 ZZZ2  | S2 = Y2 ZZZ1      (R = S2 - S1)
//   3: PP = P^2      | RR = R^2
//   4: PPP = PP P    | Q = U1 PP
//   5: ZZ1 ZZ2       | ZZZ1 ZZZ2         (X3 = RR - PPP - 2Q, one fold pass)
//   6: ZZ3 = ZZ1 ZZ2 PP | ZZZ3 = ZZZ1 ZZZ2 PPP
//   7: R (Q - X3)    | S1 PPP            (Y3 = difference of the two)
// r may alias p or q: every input is read before r is written.
template <class G>
__device__ __forceinline__ void split_add(Xyzz<G>& r, const Xyzz<G>& p, const Xyzz<G>& q) {
    using O = typename Split<G>::F;
    if (xyzz_is_inf(p)) {
        r = q;
        return;
    }
    if (xyzz_is_inf(q)) {
        r = p;
        return;
    }
    const bool h = Split<G>::half();
    Fp U1, U2, S1, S2, P, R, PP, RR, PPP, Q, T1, T2, X3, Z2, Z3, Ya, Yb, a, b, d, u;
    fp_sel(a, h, q.X, p.X);
    fp_sel(b, h, p.ZZ, q.ZZ);
    O::mul(u, a, b);
    Split<G>::results(U1, U2, u);
    fp_sel(a, h, q.Y, p.Y);
    fp_sel(b, h, p.ZZZ, q.ZZZ);
    O::mul(u, a, b);
    Split<G>::results(S1, S2, u);
    O::sub_raw(P, U2, U1);  // raw where the field has it: product operands only
    O::sub_raw(R, S2, S1);
    fp_sel(a, h, R, P);
    O::sqr_raw(u, a);
    Split<G>::results(PP, RR, u);
    if (O::is_zero(PP)) {  // P = 0: the same x
        if (O::is_zero(RR)) {
            split_dbl<G>(r, p);
        } else {
            xyzz_set_inf(r);
        }
        return;
    }
    fp_sel(b, h, U1, P);
    O::mul_raw(u, PP, b);
    Split<G>::results(PPP, Q, u);
    fp_sel(a, h, p.ZZZ, p.ZZ);
    fp_sel(b, h, q.ZZZ, q.ZZ);
    O::mul(u, a, b);
    Split<G>::results(T1, T2, u);
    fp_sub_b2c(X3, RR, PPP, Q);
    fp_sel(a, h, T2, T1);
    fp_sel(b, h, PPP, PP);
    O::mul(u, a, b);
    Split<G>::results(Z2, Z3, u);
    O::sub_raw(d, Q, X3);
    fp_sel(a, h, S1, R);
    fp_sel(b, h, PPP, d);
    O::mul_raw(u, a, b);
    Split<G>::results(Ya, Yb, u);
    r.X = X3;
    O::sub(r.Y, Ya, Yb);
    r.ZZ = Z2;
    r.ZZZ = Z3;
}

template <>
__device__ __forceinline__ void xyzz_dbl<FpPairOps>(Xyzz<FpPairOps>& r, const Xyzz<FpPairOps>& p) {
    split_dbl<FpPairOps>(r, p);
}
template <>
__device__ __forceinline__ void xyzz_add<FpPairOps>(Xyzz<FpPairOps>& r, const Xyzz<FpPairOps>& p,
                                                    const Xyzz<FpPairOps>& q) {
    split_add<FpPairOps>(r, p, q);
}
template <>
__device__ __forceinline__ void xyzz_dbl<Fp2QuadOps>(Xyzz<Fp2QuadOps>& r, const Xyzz<Fp2QuadOps>& p) {
    split_dbl<Fp2QuadOps>(r, p);
}
template <>
__device__ __forceinline__ void xyzz_add<Fp2QuadOps>(Xyzz<Fp2QuadOps>& r, const Xyzz<Fp2QuadOps>& p,
                                                     const Xyzz<Fp2QuadOps>& q) {
    split_add<Fp2QuadOps>(r, p, q);
}

}  // namespace gm
