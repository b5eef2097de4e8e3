/*
 * C++ API of Fp2 = Fp[i]/(i^2 + 1): a drop-in for the reference's
 * include/fp2.h. The host element functions (reference fp2.h:20-89) are
 * implemented in gpu_msm_amd/csrc/host_api.hip; the stream functions
 * (reference fp2.h:92-134, src/fp2_kernels.cu:564-678; cudaStream_t becomes
 * hipStream_t) in gpu_msm_amd/csrc/surface.hip.
 *
 * In the stream functions every pointer is a DEVICE pointer to one element
 * (c0, c1), 14 x u64.
 * Semantics follow reference src/fp2.cu: fp2_mul / fp2_square / fp2_inv /
 * fp2_div on normal-form values (Karatsuba, fp2.cu:58-71 and 118-129; norm inverse,
 * fp2.cu:201-214; division fp2.cu:262-266). Inputs below p, canonical results; null pointers and HIP
 * errors abort() with a message.
 */
#ifndef GPU_MSM_FP2_H
#define GPU_MSM_FP2_H

#ifndef __cplusplus
#error "fp2.h is the C++ surface"
#endif

#include <hip/hip_runtime_api.h>

#include "fp.h"
#include "gpu_msm_types.h"

/* ------------------------------------------------- host element functions
 * reference include/fp2.h:20-89 (src/fp2.cu:10-289), run on the host by this
 * library; outputs may alias inputs. fp2_mul, fp2_square, fp2_inv and
 * fp2_div take and return normal-form values, fp2_mont_mul and fp2_mont_inv
 * Montgomery images. */
int fp2_cmp(const Fp2& a, const Fp2& b);  /* c0 first, then c1 */
bool fp2_is_zero(const Fp2& a);
bool fp2_is_one(const Fp2& a);
void fp2_zero(Fp2& a);
void fp2_one(Fp2& a);  /* 1 + 0 i, normal form */
void fp2_copy(Fp2& dst, const Fp2& src);
void fp2_add(Fp2& c, const Fp2& a, const Fp2& b);
void fp2_sub(Fp2& c, const Fp2& a, const Fp2& b);
void fp2_mul(Fp2& c, const Fp2& a, const Fp2& b);
void fp2_mont_mul(Fp2& c, const Fp2& a, const Fp2& b);
void fp2_square(Fp2& c, const Fp2& a);
void fp2_neg(Fp2& c, const Fp2& a);
void fp2_conjugate(Fp2& c, const Fp2& a);
void fp2_inv(Fp2& c, const Fp2& a);       /* 0 -> 0 */
void fp2_mont_inv(Fp2& c, const Fp2& a);  /* 0 -> 0 */
void fp2_div(Fp2& c, const Fp2& a, const Fp2& b);
void fp2_cmov(Fp2& dst, const Fp2& src, uint64_t condition);
void fp2_frobenius(Fp2& c, const Fp2& a);  /* a^p = conjugate(a) */
void fp2_mul_by_i(Fp2& c, const Fp2& a);
void fp2_mul_by_non_residue(Fp2& c, const Fp2& a);  /* the non-residue is i */

/* ------------------------------------------------------- stream functions */
void fp2_add_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_add(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_sub_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_sub(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_mul_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_mul(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_square_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_square(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_neg_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_neg(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_conjugate_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_conjugate(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_inv_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_inv(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a);
void fp2_div_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_div(hipStream_t stream, uint32_t gpu_index, Fp2* d_c, const Fp2* d_a, const Fp2* d_b);
void fp2_copy_async(hipStream_t stream, uint32_t gpu_index, Fp2* d_dst, const Fp2* d_src);
void fp2_copy(hipStream_t stream, uint32_t gpu_index, Fp2* d_dst, const Fp2* d_src);

#endif /* GPU_MSM_FP2_H */
