/*
 * C ABI of the BLS12-446 MSM library — drop-in for the reference's
 * rust-api/include/rust_wrapper.h (same names, argument meaning, return codes
 * and input-form conventions; implemented in gpu_msm_amd/csrc/capi.hip).
 *
 * Every pointer is a HOST pointer owned by the caller.
 */
#ifndef GPU_MSM_RUST_WRAPPER_H
#define GPU_MSM_RUST_WRAPPER_H

#include "gpu_msm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference rust_wrapper.h:14 / c_wrapper.cu:12-14:
 * infinity -> (0,0,0), else (x, y, R mod p). */
void affine_to_projective_g1_wrapper(G1ProjectivePoint* proj, const G1Point* affine);
/* reference rust_wrapper.h:17 / c_wrapper.cu:17-19 */
void affine_to_projective_g2_wrapper(G2ProjectivePoint* proj, const G2Point* affine);

/* reference rust_wrapper.h:20 / c_wrapper.cu:22-24:
 * Z = 0 -> infinity, else (X Z^-1, Y Z^-1), canonical, Montgomery form. */
void projective_to_affine_g1_wrapper(G1Point* affine, const G1ProjectivePoint* proj);
/* reference rust_wrapper.h:23 / c_wrapper.cu:27-29 */
void projective_to_affine_g2_wrapper(G2Point* affine, const G2ProjectivePoint* proj);

/* reference rust_wrapper.h:26-35 / c_wrapper.cu:32-49 */
void g1_point_at_infinity_wrapper(G1Point* point);
void g2_point_at_infinity_wrapper(G2Point* point);
void g1_projective_point_at_infinity_wrapper(G1ProjectivePoint* point);
void g2_projective_point_at_infinity_wrapper(G2ProjectivePoint* point);

/* reference rust_wrapper.h:38-41 / c_wrapper.cu:52-59 */
bool g1_is_infinity_wrapper(const G1Point* point);
bool g2_is_infinity_wrapper(const G2Point* point);

/* reference rust_wrapper.h:45-51 / c_wrapper.cu:64-172.
 * result = sum_i scalars[i] * points[i], one u64 scalar per point.
 * G1 input points are in STANDARD form (converted to Montgomery inside, as
 * reference c_wrapper.cu:124-127 does). Output: projective, Montgomery form.
 * Returns 0 ok; -1 n <= 0; -2 device/stream setup failed; -3 allocation
 * failed; -4 HIP error during conversion or MSM. */
int g1_msm_wrapper(G1ProjectivePoint* result, const G1Point* points, const uint64_t* scalars, int n,
                   uint32_t gpu_index);

/* reference rust_wrapper.h:53-59 / c_wrapper.cu:175-258.
 * G2 input points are taken as already in MONTGOMERY form (no conversion, as
 * reference c_wrapper.cu:223-224). Same output form and return codes. */
int g2_msm_wrapper(G2ProjectivePoint* result, const G2Point* points, const uint64_t* scalars, int n,
                   uint32_t gpu_index);

/* reference rust-api/src/ffi.rs:67 binds init_device_generators(stream,
 * gpu_index) with C linkage (reference src/curve.cu:833-838). Generators are
 * compile-time constants here, so this only selects and checks the device;
 * `stream` is ignored and may be NULL. */
void init_device_generators(const void* stream, uint32_t gpu_index);

/* reference rust_wrapper.h:62-64 / c_wrapper.cu:261-288 */
void g1_from_montgomery_wrapper(G1Point* result, const G1Point* point);
void g2_from_montgomery_wrapper(G2Point* result, const G2Point* point);
void fp_to_montgomery_wrapper(Fp* result, const Fp* value);

#ifdef __cplusplus
}
#endif

#endif /* GPU_MSM_RUST_WRAPPER_H */
