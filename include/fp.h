/*
 * C++ API of the base field Fp: a drop-in for the reference's include/fp.h.
 * The host element functions (reference fp.h:34-148) are implemented in
 * gpu_msm_amd/csrc/host_api.hip; the stream functions (reference
 * fp.h:156-205, src/fp_kernels.cu:753-906; cudaStream_t becomes hipStream_t)
 * in gpu_msm_amd/csrc/surface.hip.
 *
 * In the stream functions every pointer is a DEVICE pointer to one 7 x u64
 * little-endian element.
 * `_async` functions enqueue on `stream`; the others also synchronise it.
 * Semantics follow reference src/fp.cu: fp_mul / fp_inv / fp_div work on
 * normal-form values (fp.cu:424-433, 522-582), fp_mont_mul on Montgomery
 * images; inputs below p, canonical results. Null pointers and HIP errors
 * print a message and abort(), as the reference's PANIC does.
 */
#ifndef GPU_MSM_FP_H
#define GPU_MSM_FP_H

#ifndef __cplusplus
#error "fp.h is the C++ surface"
#endif

#include <hip/hip_runtime_api.h>

#include "gpu_msm_types.h"

/* ------------------------------------------------- host element functions
 * reference include/fp.h:34-148 (src/fp.cu:42-720), run on the host by this
 * library. Values are 7 x u64 little-endian; outputs may alias inputs.
 * fp_mul, fp_inv, fp_div, fp_pow, fp_pow_u64, fp_sqrt and
 * fp_is_quadratic_residue take and return normal-form values; fp_mont_mul,
 * fp_mont_inv and fp_mont_reduce work on Montgomery images (R = 2^448).
 * Inputs below p give canonical results. */
int fp_cmp(const Fp& a, const Fp& b);  /* -1, 0, 1 on the limbs */
bool fp_is_zero(const Fp& a);
bool fp_is_one(const Fp& a);           /* a == 1 (normal form) */
void fp_zero(Fp& a);
void fp_one(Fp& a);                    /* 1 in normal form */
void fp_one_montgomery(Fp& a);         /* R mod p */
void fp_copy(Fp& dst, const Fp& src);
uint64_t fp_add_raw(Fp& c, const Fp& a, const Fp& b);  /* 448-bit sum, returns the carry */
uint64_t fp_sub_raw(Fp& c, const Fp& a, const Fp& b);  /* 448-bit difference, returns the borrow */
void fp_add(Fp& c, const Fp& a, const Fp& b);
void fp_sub(Fp& c, const Fp& a, const Fp& b);
void fp_mul_raw(uint64_t* c, const Fp& a, const Fp& b);  /* 14-word product */
void fp_reduce(Fp& c, const uint64_t* a);       /* = fp_mont_reduce, as in the reference */
void fp_mont_reduce(Fp& c, const uint64_t* a);  /* a * R^-1 mod p for a 14-word a */
void fp_mont_mul(Fp& c, const Fp& a, const Fp& b);
void fp_to_montgomery(Fp& c, const Fp& a);
void fp_from_montgomery(Fp& c, const Fp& a);
void fp_to_montgomery_batch(Fp* dst, const Fp* src, int n);    /* host arrays */
void fp_from_montgomery_batch(Fp* dst, const Fp* src, int n);  /* host arrays */
void fp_mul(Fp& c, const Fp& a, const Fp& b);
void fp_neg(Fp& c, const Fp& a);
void fp_inv(Fp& c, const Fp& a);       /* 0 -> 0 */
void fp_mont_inv(Fp& c, const Fp& a);  /* 0 -> 0 */
void fp_div(Fp& c, const Fp& a, const Fp& b);  /* b = 0 -> 0 */
void fp_pow(Fp& c, const Fp& a, const uint64_t* e, int e_limbs);  /* at most 7 exponent words are read */
void fp_pow_u64(Fp& c, const Fp& a, uint64_t e);
bool fp_sqrt(Fp& c, const Fp& a);      /* a^((p+1)/4); false and c = 0 for a non-residue */
bool fp_is_quadratic_residue(const Fp& a);
void fp_cmov(Fp& dst, const Fp& src, uint64_t condition);  /* moves when bit 0 of condition is set */
const Fp& fp_modulus();

/* ------------------------------------------------------- stream functions */
void fp_add_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_add(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_sub_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_sub(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_mul_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_mul(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_neg_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_neg(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_inv_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_inv(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_div_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_div(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_copy_async(hipStream_t stream, uint32_t gpu_index, Fp* d_dst, const Fp* d_src);
void fp_copy(hipStream_t stream, uint32_t gpu_index, Fp* d_dst, const Fp* d_src);
void fp_zero_async(hipStream_t stream, uint32_t gpu_index, Fp* d_a);
void fp_zero(hipStream_t stream, uint32_t gpu_index, Fp* d_a);
void fp_one_async(hipStream_t stream, uint32_t gpu_index, Fp* d_a);  /* 1 in normal form */
void fp_one(hipStream_t stream, uint32_t gpu_index, Fp* d_a);
void fp_to_montgomery_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_to_montgomery(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_from_montgomery_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_from_montgomery(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a);
void fp_mont_mul_async(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);
void fp_mont_mul(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b);

#endif /* GPU_MSM_FP_H */
