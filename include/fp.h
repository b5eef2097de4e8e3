/*
 * C++ device-pointer API of the base field Fp: a drop-in for the stream
 * functions of the reference's include/fp.h:156-205 (implemented in
 * reference src/fp_kernels.cu:753-906). cudaStream_t becomes hipStream_t.
 * Implemented in gpu_msm_amd/csrc/surface.hip.
 *
 * Every pointer is a DEVICE pointer to one 7 x u64 little-endian element.
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
