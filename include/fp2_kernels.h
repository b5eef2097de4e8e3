/*
 * Element-wise Fp2 kernels and host-pointer single operations: a drop-in for
 * the reference's include/fp2_kernels.h:1-40 (src/fp2_kernels.cu:115-560).
 * cudaStream_t becomes hipStream_t. Implemented in gpu_msm_amd/csrc/surface.hip.
 *
 * *_array_host: HOST pointers, n elements, synchronous on `stream`.
 * The stream-less *_gpu functions use the current device and its null stream
 * (as reference src/fp2_kernels.cu:218-240 does) and are synchronous.
 */
#ifndef GPU_MSM_FP2_KERNELS_H
#define GPU_MSM_FP2_KERNELS_H

#ifndef __cplusplus
#error "fp2_kernels.h is the C++ surface"
#endif

#include <hip/hip_runtime_api.h>

#include "gpu_msm_types.h"

void fp2_add_array_host(hipStream_t stream, uint32_t gpu_index, Fp2* c, const Fp2* a, const Fp2* b, int n);
void fp2_mul_array_host(hipStream_t stream, uint32_t gpu_index, Fp2* c, const Fp2* a, const Fp2* b, int n);

void fp2_add_gpu(Fp2* result, const Fp2* a, const Fp2* b);
void fp2_sub_gpu(Fp2* result, const Fp2* a, const Fp2* b);
void fp2_mul_gpu(Fp2* result, const Fp2* a, const Fp2* b);
void fp2_neg_gpu(Fp2* result, const Fp2* a);
void fp2_conjugate_gpu(Fp2* result, const Fp2* a);
void fp2_square_gpu(Fp2* result, const Fp2* a);
void fp2_inv_gpu(Fp2* result, const Fp2* a);
void fp2_div_gpu(Fp2* result, const Fp2* a, const Fp2* b);
void fp2_mul_by_i_gpu(Fp2* result, const Fp2* a);
void fp2_frobenius_gpu(Fp2* result, const Fp2* a); /* conjugation, reference src/fp2.cu:277-279 */
int fp2_cmp_gpu(const Fp2* a, const Fp2* b);
bool fp2_is_zero_gpu(const Fp2* a);
bool fp2_is_one_gpu(const Fp2* a);
void fp2_copy_gpu(Fp2* result, const Fp2* a);
void fp2_cmov_gpu(Fp2* result, const Fp2* src, uint64_t condition);

#endif /* GPU_MSM_FP2_KERNELS_H */
