/*
 * C++ device-pointer API of Fp2 = Fp[i]/(i^2 + 1): a drop-in for the stream
 * functions of the reference's include/fp2.h:92-134 (implemented in
 * reference src/fp2_kernels.cu:564-678). cudaStream_t becomes hipStream_t.
 * Implemented in gpu_msm_amd/csrc/surface.hip.
 *
 * Every pointer is a DEVICE pointer to one element (c0, c1), 14 x u64.
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

#include "gpu_msm_types.h"

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
