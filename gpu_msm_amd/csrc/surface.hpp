// Host-side interface of the single-operation kernels (point operations,
// Montgomery batch conversion, Fp / Fp2 element operations) and of the
// descriptor upload used by the device-pointer MSM entry points.
#pragma once
#include "engine.hpp"

namespace gm {

enum PointOp : int { PO_ADD = 0, PO_DOUBLE, PO_NEG, PO_INF, PO_SMUL, PO_TO_MONT, PO_FROM_MONT };

enum FieldOp : int {
    FO_ADD = 0,
    FO_SUB,
    FO_MUL,        // a b mod p, normal form (reference src/fp.cu:424-433)
    FO_MONT_MUL,   // a b R^-1 mod p
    FO_SQUARE,     // Fp2 only, normal form
    FO_NEG,
    FO_CONJ,       // Fp2 only
    FO_MUL_BY_I,   // Fp2 only
    FO_INV,        // normal form, 0 -> 0
    FO_DIV,        // a / b, normal form, b = 0 -> 0
    FO_TO_MONT,
    FO_FROM_MONT,
    FO_COPY,
    FO_ZERO,
    FO_ONE,        // 1 in normal form
    FO_CMOV,       // c = cond & 1 ? a : c
    FO_SQRT,       // Fp only: a^((p+1)/4) if a is a square (flag 1), else 0 (flag 0)
    FO_IS_QR,      // Fp only: Euler's criterion, 0 counts as a square
    FO_POW_U64,    // Fp only: a^e, normal form
    FO_CMP,        // raw limb comparison, -1 / 0 / 1
    FO_IS_ZERO,    // raw
    FO_IS_ONE,     // raw normal-form one
    FO_COUNT
};

template <class O>
hipError_t launch_point_op(hipStream_t s, int op, void* r, const void* a, const void* b, const uint64_t* d_sc,
                           uint64_t sc, int limbs);
template <class O>
hipError_t launch_point_mont_batch(hipStream_t s, void* pts, int n, bool to_mont);
template <class O>
hipError_t launch_field_op(hipStream_t s, int op, void* c, const void* a, const void* b, int n, uint64_t arg,
                           int* iout);
// Writes one MSM descriptor into device memory from a kernel argument, so the
// device-pointer MSM needs no host staging buffer and stays stream-ordered.
hipError_t launch_put_desc(hipStream_t s, MsmDesc* dst, const MsmDesc& d);

}  // namespace gm
