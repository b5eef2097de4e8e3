/*
 * C++ API of the curve layer: a drop-in for the reference's include/curve.h
 * (same names, argument meaning and PointType instantiations; cudaStream_t
 * becomes hipStream_t). The host functions (reference curve.h:50-134) are in
 * gpu_msm_amd/csrc/host_api.hip; the device-pointer functions (reference
 * curve.h:140-241) in gpu_msm_amd/csrc/surface.hip on the MI355X kernels of
 * this library.
 *
 * In the device-pointer functions every pointer is a DEVICE pointer. `_async` functions enqueue on `stream`
 * and return; the others also synchronise `stream`. PointType is G1Point or
 * G2Point: affine records in Montgomery form with field values below p.
 * Results are canonical, so they equal the reference's bit for bit.
 *
 * Errors behave as the reference's PANIC / check_cuda_error
 * (reference include/device.h:11-39): a null pointer, a negative size or a HIP
 * error prints a message naming the function and calls abort().
 */
#ifndef GPU_MSM_CURVE_H
#define GPU_MSM_CURVE_H

#ifndef __cplusplus
#error "curve.h is the C++ surface; C callers use rust_wrapper.h and gpu_msm_ext.h"
#endif

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "fp.h"
#include "fp2.h"
#include "gpu_msm_types.h"

/* ------------------------------------------------- host curve functions
 * reference include/curve.h:50-134 (src/curve.cu:14-778), run on the host by
 * this library (gpu_msm_amd/csrc/host_api.hip). Coordinates are Montgomery
 * images except where a function says otherwise; outputs may alias inputs.
 *
 * curve_b_g1 / curve_b_g2 return the curves' real coefficients in normal
 * form, b = 1 and b' = 1 + i, and is_on_curve_g1 / _g2 (normal-form points,
 * as the reference's fp_mul-based check) test against them: the reference's
 * constants 4 and 4i (src/curve.cu:280-289) reject its own generators. */
const Fp& curve_b_g1();
const Fp2& curve_b_g2();
bool is_on_curve_g1(const G1Point& point);
bool is_on_curve_g2(const G2Point& point);
void g1_point_at_infinity(G1Point& point);  /* (0, 0, infinity) */
void g2_point_at_infinity(G2Point& point);
bool g1_is_infinity(const G1Point& point);
bool g2_is_infinity(const G2Point& point);

/* Affine group law (reference src/curve.cu:72-268): O + P copies P as stored,
 * P + (-P) and points with y = 0 doubled give O. point_scalar_mul reads
 * scalar_limbs little-endian u64 words. Instantiated for G1Point and G2Point. */
template <typename PointType>
void point_add(PointType& result, const PointType& p1, const PointType& p2);
template <typename PointType>
void point_double(PointType& result, const PointType& p);
template <typename PointType>
void point_neg(PointType& result, const PointType& p);
template <typename PointType>
void point_scalar_mul(PointType& result, const PointType& point, const uint64_t* scalar, int scalar_limbs);

/* Generators in STANDARD form (reference src/curve.cu:840-875 on the host);
 * callers convert with fp_to_montgomery. */
const G1Point& g1_generator();
const G2Point& g2_generator();

/* The device-side generator set-up of the reference (src/curve.cu:816-838);
 * the same C symbol as rust_wrapper.h declares. */
extern "C" void init_device_generators(const void* stream, uint32_t gpu_index);

void g1_projective_point_at_infinity(G1ProjectivePoint& p);  /* all zero */
void g2_projective_point_at_infinity(G2ProjectivePoint& p);
/* (x, y) -> (x, y, R mod p); O -> all zero */
void affine_to_projective(G1ProjectivePoint& proj, const G1Point& affine);
void affine_to_projective(G2ProjectivePoint& proj, const G2Point& affine);
/* (X / Z, Y / Z); Z = 0 -> O */
void projective_to_affine_g1(G1Point& affine, const G1ProjectivePoint& proj);
void projective_to_affine_g2(G2Point& affine, const G2ProjectivePoint& proj);
/* The reference's standard-projective formulas (src/curve.cu:444-738), so the
 * raw (X:Y:Z) matches it, not only the affine image; Z = 0 inputs return the
 * other operand. */
void projective_point_add(G1ProjectivePoint& result, const G1ProjectivePoint& p1, const G1ProjectivePoint& p2);
void projective_point_add(G2ProjectivePoint& result, const G2ProjectivePoint& p1, const G2ProjectivePoint& p2);
void projective_point_double(G1ProjectivePoint& result, const G1ProjectivePoint& p);
void projective_point_double(G2ProjectivePoint& result, const G2ProjectivePoint& p);

/* ------------------------------------------- single-point operations
 * reference include/curve.h:150-206, src/curve.cu:2386-2525; formulas
 * src/curve.cu:72-268 (affine add / double / negate / double-and-add). */

/* d_result = d_p1 + d_p2; O + P copies P as stored (reference curve.cu:178-189). */
template <typename PointType>
void point_add_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_p1,
                     const PointType* d_p2);
template <typename PointType>
void point_add(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_p1,
               const PointType* d_p2);

/* d_result = 2 d_p; O and points with y = 0 give O. */
template <typename PointType>
void point_double_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_p);
template <typename PointType>
void point_double(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_p);

/* d_result = -d_p (x kept, y -> p - y). */
template <typename PointType>
void point_neg_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_p);
template <typename PointType>
void point_neg(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_p);

/* d_result = scalar * d_point (one u64 scalar). */
template <typename PointType>
void point_scalar_mul_u64_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result,
                                const PointType* d_point, uint64_t scalar);
template <typename PointType>
void point_scalar_mul_u64(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_point,
                          uint64_t scalar);

/* d_result = scalar * d_point, scalar = scalar_limbs little-endian u64 words at d_scalar (device). */
template <typename PointType>
void point_scalar_mul_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_point,
                            const uint64_t* d_scalar, int scalar_limbs);
template <typename PointType>
void point_scalar_mul(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_point,
                      const uint64_t* d_scalar, int scalar_limbs);

/* d_result = O: x = y = 0, infinity = true. */
template <typename PointType>
void point_at_infinity_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result);
template <typename PointType>
void point_at_infinity(hipStream_t stream, uint32_t gpu_index, PointType* d_result);

/* Coordinates times R (to) or R^-1 (from) mod p; O gives O. */
template <typename PointType>
void point_to_montgomery_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_point);
template <typename PointType>
void point_to_montgomery(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_point);
template <typename PointType>
void point_from_montgomery_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result,
                                 const PointType* d_point);
template <typename PointType>
void point_from_montgomery(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_point);

/* In place over n records; records at infinity are left unchanged
 * (reference curve.cu:2333-2352, 2506-2522). */
template <typename PointType>
void point_to_montgomery_batch_async(hipStream_t stream, uint32_t gpu_index, PointType* d_points, int n);
template <typename PointType>
void point_to_montgomery_batch(hipStream_t stream, uint32_t gpu_index, PointType* d_points, int n);

/* ------------------------------------------------------------------ MSM
 * reference include/curve.h:212-241, src/curve.cu:2538-2826.
 *
 * d_points: n affine bases in Montgomery form. d_result: projective
 * (x, y, R mod p) or all-zero for O (u64 functions), or the canonical affine
 * record (PointType functions). n = 0 gives O; n < 0 aborts.
 *
 * Scratch: the reference sizes d_scratch as (blocks + 1) * 16 points; this
 * library's pipeline (sorted bucket lists, per-window buckets) needs the byte
 * count returned by the *_scratch_bytes query for the same n (and limbs).
 * A caller that cannot size it passes d_scratch = NULL: the call then takes
 * and returns the memory stream-ordered (hipMallocAsync / hipFreeAsync). A
 * non-NULL d_scratch is used only when the bytes from it to the end of its
 * device allocation (hipMemGetAddressRange) cover that count; a shorter
 * buffer, such as one sized the reference's way, is left untouched and the
 * call takes stream-ordered memory as for NULL.
 *
 * The _async forms never block the host: the window sums go to pinned host
 * memory, a host function enqueued on the stream (hipLaunchHostFunc) runs the
 * window combine, and the record is copied to d_result, all in stream order
 * (a serial chain of ~bits doublings runs about ten times faster on a host
 * core than on one GPU lane). The synchronous forms do the same combine
 * inline and return when d_result is written. Both give identical results.
 * Calls on different streams may overlap when their scratch buffers differ.
 * The _async forms are not meant for hipGraph capture (their pinned slot is
 * taken when the call is made). */
/* The reference's legacy sizing constants (reference include/curve.h:143-145),
 * kept so callers that size scratch with them still compile; such a buffer is
 * too small for this pipeline and the MSM then takes its own (above). */
#define MSM_WINDOW_SIZE 4
#define MSM_BUCKET_COUNT 16
#define MSM_SIGNED_BUCKET_COUNT 8

size_t point_msm_u64_scratch_bytes_g1(int n);
size_t point_msm_u64_scratch_bytes_g2(int n);
template <typename PointType>
size_t point_msm_scratch_bytes(int n, int scalar_limbs);

void point_msm_u64_async_g1(hipStream_t stream, uint32_t gpu_index, G1ProjectivePoint* d_result,
                            const G1Point* d_points, const uint64_t* d_scalars, G1ProjectivePoint* d_scratch, int n);
void point_msm_u64_g1(hipStream_t stream, uint32_t gpu_index, G1ProjectivePoint* d_result, const G1Point* d_points,
                      const uint64_t* d_scalars, G1ProjectivePoint* d_scratch, int n);
void point_msm_u64_async_g2(hipStream_t stream, uint32_t gpu_index, G2ProjectivePoint* d_result,
                            const G2Point* d_points, const uint64_t* d_scalars, G2ProjectivePoint* d_scratch, int n);
void point_msm_u64_g2(hipStream_t stream, uint32_t gpu_index, G2ProjectivePoint* d_result, const G2Point* d_points,
                      const uint64_t* d_scalars, G2ProjectivePoint* d_scratch, int n);

/* Multi-limb scalars: scalar_limbs (1..8) little-endian u64 words per point,
 * any value below 2^(64 scalar_limbs); affine result (reference
 * src/curve.cu:2763-2826). */
template <typename PointType>
void point_msm_async(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_points,
                     const uint64_t* d_scalars, int scalar_limbs, PointType* d_scratch, int n);
template <typename PointType>
void point_msm(hipStream_t stream, uint32_t gpu_index, PointType* d_result, const PointType* d_points,
               const uint64_t* d_scalars, int scalar_limbs, PointType* d_scratch, int n);

#endif /* GPU_MSM_CURVE_H */
