// XYZZ addition and doubling for FpPairOps: the two lanes of a pair hold the
// same operands and compute two independent field products per step, so a
// point operation's serial depth is about half its product count. Used by
// the G1 bucket-reduction and window-sum kernels, which run at under one wave
// per SIMD, where a wave's instruction issue (not the SIMD's throughput)
// bounds their time. Representations may differ from the one-lane formulas
// (X3 through normalised subtractions instead of the fused square); every
// value is the same point, and results are compared only in affine form.
#pragma once
#include "curve.hpp"

namespace gm {

// r = 2p (dbl-2008-s-1), four paired steps:
//   1: V = U^2 | T = X^2            (U = 2Y)
//   2: W = U V | S = X V            (M = 3T)
//   3: M^2     | ZZ3 = V ZZ         (X3 = M^2 - 2S)
//   4: Y3 = M (S - X3) - W Y | ZZZ3 = W ZZZ   (both as a difference of products)
template <>
__device__ __forceinline__ void xyzz_dbl<FpPairOps>(Xyzz<FpPairOps>& r, const Xyzz<FpPairOps>& p) {
    if (xyzz_is_inf(p)) {
        r = p;
        return;
    }
    const bool q = lane_part() != 0;
    Fp U, V, T, W, S, M, MM, X3, Z2, Y3, Z3, a, b, c, d, u;
    fp_dbl(U, p.Y);
    fp_sel(a, q, p.X, U);
    fp_sqr(u, a);
    pair_results(V, T, u);
    fp_sel(a, q, p.X, U);
    fp_mul(u, a, V);
    pair_results(W, S, u);
    fp_dbl(M, T);
    fp_add(M, M, T);
    fp_sel(a, q, V, M);
    fp_sel(b, q, p.ZZ, M);
    fp_mul(u, a, b);
    pair_results(MM, Z2, u);
    fp_sub(X3, MM, S);
    fp_sub(X3, X3, S);
    fp_sub_raw(d, S, X3);  // raw: a product operand only
    Fp zero;
    fp_set_zero(zero);
    fp_sel(a, q, W, M);
    fp_sel(b, q, p.ZZZ, d);
    fp_sel(c, q, zero, W);
    fp_sel(d, q, zero, p.Y);
    fp_mul_sub_raw(u, a, b, c, d);  // part 0: M (S - X3) - W Y; part 1: W ZZZ - 0
    pair_results(Y3, Z3, u);
    r.X = X3;
    r.Y = Y3;
    r.ZZ = Z2;
    r.ZZZ = Z3;
}

// r = p + q (add-2008-s), seven paired steps:
//   1: U1 = X1 ZZ2   | U2 = X2 ZZ1       (P = U2 - U1, raw)
//   2: S1 = Y1 ZZZ2  | S2 = Y2 ZZZ1      (R = S2 - S1, raw)
//   3: PP = P^2      | RR = R^2
//   4: PPP = PP P    | Q = U1 PP
//   5: ZZ1 ZZ2       | ZZZ1 ZZZ2         (X3 = RR - PPP - 2Q)
//   6: ZZ3 = ZZ1 ZZ2 PP | ZZZ3 = ZZZ1 ZZZ2 PPP
//   7: R (Q - X3)    | S1 PPP            (Y3 = difference of the two)
// r may alias p or q: every input is read before r is written.
template <>
__device__ __forceinline__ void xyzz_add<FpPairOps>(Xyzz<FpPairOps>& r, const Xyzz<FpPairOps>& p,
                                                    const Xyzz<FpPairOps>& q) {
    if (xyzz_is_inf(p)) {
        r = q;
        return;
    }
    if (xyzz_is_inf(q)) {
        r = p;
        return;
    }
    const bool h = lane_part() != 0;
    Fp U1, U2, S1, S2, P, R, PP, RR, PPP, Q, T1, T2, X3, Z2, Z3, Ya, Yb, a, b, u;
    fp_sel(a, h, q.X, p.X);
    fp_sel(b, h, p.ZZ, q.ZZ);
    fp_mul(u, a, b);
    pair_results(U1, U2, u);
    fp_sel(a, h, q.Y, p.Y);
    fp_sel(b, h, p.ZZZ, q.ZZZ);
    fp_mul(u, a, b);
    pair_results(S1, S2, u);
    fp_sub_raw(P, U2, U1);
    fp_sub_raw(R, S2, S1);
    fp_sel(a, h, R, P);
    fp_sqr_raw(u, a);
    pair_results(PP, RR, u);
    if (fp_is_zero(PP)) {  // P = 0 (mod p): the same x
        if (fp_is_zero(RR)) {
            xyzz_dbl<FpPairOps>(r, p);
        } else {
            xyzz_set_inf(r);
        }
        return;
    }
    fp_sel(b, h, U1, P);
    fp_mul_raw(u, PP, b);
    pair_results(PPP, Q, u);
    fp_sel(a, h, p.ZZZ, p.ZZ);
    fp_sel(b, h, q.ZZZ, q.ZZ);
    fp_mul(u, a, b);
    pair_results(T1, T2, u);
    fp_sub(X3, RR, PPP);
    fp_sub(X3, X3, Q);
    fp_sub(X3, X3, Q);
    fp_sel(a, h, T2, T1);
    fp_sel(b, h, PPP, PP);
    fp_mul(u, a, b);
    pair_results(Z2, Z3, u);
    Fp d;
    fp_sub_raw(d, Q, X3);  // raw: a product operand only
    fp_sel(a, h, S1, R);
    fp_sel(b, h, PPP, d);
    fp_mul_raw(u, a, b);
    pair_results(Ya, Yb, u);
    r.X = X3;
    fp_sub(r.Y, Ya, Yb);
    r.ZZ = Z2;
    r.ZZZ = Z3;
}

}  // namespace gm
