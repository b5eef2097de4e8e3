/*
 * Element-wise Fp kernels and host-pointer single operations: a drop-in for
 * the reference's include/fp_kernels.h:1-64 (src/fp_kernels.cu:151-751).
 * cudaStream_t becomes hipStream_t. Implemented in gpu_msm_amd/csrc/surface.hip.
 *
 * *_array_device: DEVICE pointers, n elements, enqueued on `stream` (no sync).
 * *_array_host and *_gpu: HOST pointers; device memory is taken and returned
 * stream-ordered and the call synchronises `stream` before returning.
 * fp_mul* multiply normal-form values (reference src/fp.cu:424-433).
 * n < 0 and null pointers abort() with a message; n = 0 does nothing.
 */
#ifndef GPU_MSM_FP_KERNELS_H
#define GPU_MSM_FP_KERNELS_H

#ifndef __cplusplus
#error "fp_kernels.h is the C++ surface"
#endif

#include <hip/hip_runtime_api.h>

#include "gpu_msm_types.h"

void fp_add_array_host(hipStream_t stream, uint32_t gpu_index, Fp* c, const Fp* a, const Fp* b, int n);
void fp_mul_array_host(hipStream_t stream, uint32_t gpu_index, Fp* c, const Fp* a, const Fp* b, int n);
void fp_add_array_device(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b, int n);
void fp_mul_array_device(hipStream_t stream, uint32_t gpu_index, Fp* d_c, const Fp* d_a, const Fp* d_b, int n);

void fp_add_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a, const Fp* b);
void fp_sub_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a, const Fp* b);
void fp_mul_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a, const Fp* b);
void fp_neg_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a);
void fp_inv_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a);
void fp_div_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a, const Fp* b);
int fp_cmp_gpu(hipStream_t stream, uint32_t gpu_index, const Fp* a, const Fp* b);
bool fp_is_zero_gpu(hipStream_t stream, uint32_t gpu_index, const Fp* a);
bool fp_is_one_gpu(hipStream_t stream, uint32_t gpu_index, const Fp* a);
void fp_to_montgomery_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a);
void fp_from_montgomery_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a);
void fp_mont_mul_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a, const Fp* b);
void fp_copy_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a);
void fp_cmov_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* src, uint64_t condition);
/* a^((p+1)/4) if a is a square (true), else 0 (false); reference src/fp.cu:632-701 */
bool fp_sqrt_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* a);
bool fp_is_quadratic_residue_gpu(hipStream_t stream, uint32_t gpu_index, const Fp* a);
void fp_pow_u64_gpu(hipStream_t stream, uint32_t gpu_index, Fp* result, const Fp* base, uint64_t exp);

#endif /* GPU_MSM_FP_KERNELS_H */
