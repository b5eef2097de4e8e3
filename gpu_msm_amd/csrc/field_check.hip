// Element-wise check of the lazily reduced field products of the bucket
// kernels: G1's Fp products (FpOps, and FpAccOps with single-chain column
// sums, one element per lane) and G2's lane-pair Fp2 products (Fp2LaneOps and
// Fp2LaneAccOps, one element per lane pair). Operands are given as 28-bit
// limbs in Montgomery form, so a test can drive every product at the edges of
// its input range (values up to 2p, raw differences of magnitude up to 2p) and
// check the result (its value mod p, and that it lies in [0, 2p) with
// normalised limbs) against big integers.
#include <hip/hip_runtime.h>

#include "../../include/gpu_msm_ext.h"
#include "field.hpp"

namespace gm {

// One operation of the field policy L on operands loaded from `o`.
template <class L>
__device__ __forceinline__ void check_op(int op, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                         const uint32_t* __restrict__ c, const uint32_t* __restrict__ d,
                                         uint32_t* __restrict__ out, size_t o) {
    Fp A, B, C, D, R;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        A.v[i] = a[o + i];
        B.v[i] = b[o + i];
        C.v[i] = c[o + i];
        D.v[i] = d[o + i];
    }
    switch (op) {
        case 0: L::mul(R, A, B); break;                // A, B in [0, 2p)
        case 1: L::mul_raw(R, A, B); break;            // A, B raw
        case 2: L::mul_sub(R, A, B, C, D); break;      // all in [0, 2p)
        case 3: L::mul_sub_raw(R, A, B, C, D); break;  // A, B raw; C, D in [0, 2p)
        case 4: L::sqr_raw(R, A); break;               // A raw
        default: L::sqr_sub2(R, A, B, C); break;       // A raw; B, C in [0, 2p)
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) out[o + i] = R.v[i];
}

// G1: element t at limbs [t * 16, (t + 1) * 16).
template <class L>
__global__ void __launch_bounds__(64) k_fp_check(int op, const uint32_t* __restrict__ a,
                                                 const uint32_t* __restrict__ b, const uint32_t* __restrict__ c,
                                                 const uint32_t* __restrict__ d, uint32_t* __restrict__ out, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    check_op<L>(op, a, b, c, d, out, (size_t)t * NL);
}

// G2: element g occupies lanes 2g (c0) and 2g + 1 (c1); its component c is at
// limbs [(2g + c) * 16, (2g + c + 1) * 16) of every array.
template <class L>
__global__ void __launch_bounds__(64) k_lane_check(int op, const uint32_t* __restrict__ a,
                                                   const uint32_t* __restrict__ b, const uint32_t* __restrict__ c,
                                                   const uint32_t* __restrict__ d, uint32_t* __restrict__ out,
                                                   int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if ((t >> 1) >= n) return;  // both lanes of a pair leave together
    check_op<L>(op, a, b, c, d, out, (size_t)t * NL);
}

}  // namespace gm

extern "C" int gm_field_ops_check(int curve, int chain, int op, const void* a, const void* b, const void* c,
                                  const void* d, void* out, int n, void* stream) {
    if (n <= 0) return 0;
    if ((curve != GM_CURVE_G1 && curve != GM_CURVE_G2) || op < 0 || op > 5 || !a || !b || !c || !d || !out)
        return (int)hipErrorInvalidValue;
    const auto s = static_cast<hipStream_t>(stream);
    const auto* pa = static_cast<const uint32_t*>(a);
    const auto* pb = static_cast<const uint32_t*>(b);
    const auto* pc = static_cast<const uint32_t*>(c);
    const auto* pd = static_cast<const uint32_t*>(d);
    auto* po = static_cast<uint32_t*>(out);
    const size_t lanes = curve == GM_CURVE_G1 ? (size_t)n : 2 * (size_t)n;
    const unsigned blocks = (unsigned)((lanes + 63) / 64);
    if (curve == GM_CURVE_G1) {
        if (chain)
            gm::k_fp_check<gm::FpAccOps><<<blocks, 64, 0, s>>>(op, pa, pb, pc, pd, po, n);
        else
            gm::k_fp_check<gm::FpOps><<<blocks, 64, 0, s>>>(op, pa, pb, pc, pd, po, n);
    } else {
        if (chain)
            gm::k_lane_check<gm::Fp2LaneAccOps><<<blocks, 64, 0, s>>>(op, pa, pb, pc, pd, po, n);
        else
            gm::k_lane_check<gm::Fp2LaneOps><<<blocks, 64, 0, s>>>(op, pa, pb, pc, pd, po, n);
    }
    return (int)hipGetLastError();
}
