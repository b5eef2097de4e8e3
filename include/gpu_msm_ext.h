/*
 * Extensions of the BLS12-446 MSM C ABI beyond the reference's
 * rust_wrapper.h. Each has no direct counterpart in the reference unless its
 * comment cites one. Return codes follow the reference wrappers
 * (rust-api/src/c_wrapper.cu:71-153): 0 ok, -1 bad size/argument, -2 device
 * setup failed, -3 allocation failed, -4 HIP error; plus -5 scalar does not
 * fit the declared bit bound.
 */
#ifndef GPU_MSM_EXT_H
#define GPU_MSM_EXT_H

#include <stddef.h>

#include "gpu_msm_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Input-form and scalar flags for the *_msm_multi_wrapper entry points. */
#define GM_POINTS_MONTGOMERY 0x1u  /* points already in Montgomery form        */
#define GM_SCALARS_LT_R      0x2u  /* every scalar < r (299 bits): fewer windows */

#define GM_CURVE_G1 1
#define GM_CURVE_G2 2

/* Multi-limb MSM with host pointers: the semantics of the reference's
 * point_msm<T> (src/curve.cu:2763-2826, scalars at stride scalar_limbs
 * little-endian u64 limbs), exported with an explicit input form so G1 and G2
 * callers are symmetric. scalar_limbs in [1, 8]. Output projective,
 * Montgomery form, Z = R mod p or the all-zero point. */
int g1_msm_multi_wrapper(G1ProjectivePoint* result, const G1Point* points, const uint64_t* scalars,
                         int scalar_limbs, int n, uint32_t gpu_index, uint32_t flags);
int g2_msm_multi_wrapper(G2ProjectivePoint* result, const G2Point* points, const uint64_t* scalars,
                         int scalar_limbs, int n, uint32_t gpu_index, uint32_t flags);

/* ---------------------------------------------------------------- plans */

/* Window width and workspace the engine would use (window <= 0: heuristic). */
int gm_plan_query(int curve, int n, int scalar_limbs, int scalar_bits, int window, int* c_out, int* windows_out,
                  size_t* workspace_bytes_out);

/* ------------------------------------------------- resident device data */

typedef struct gm_bases gm_bases;     /* affine bases resident in HBM      */
typedef struct gm_scalars gm_scalars; /* scalars resident in HBM           */

/* Upload ABI points (G1Point / G2Point records), converting to the engine's
 * resident format. montgomery != 0: points already in Montgomery form. */
int gm_bases_upload(gm_bases** out, int curve, const void* host_points, int n, int montgomery, uint32_t gpu_index);
/* Bases P_i = k_i * G computed on the device from 64-bit k_i (host array). */
int gm_bases_from_k(gm_bases** out, int curve, const uint64_t* host_k, int n, uint32_t gpu_index);
/* Copy bases back as ABI records (Montgomery form). */
int gm_bases_download(const gm_bases* b, void* host_points, int offset, int n);
int gm_bases_count(const gm_bases* b);
void gm_bases_free(gm_bases* b);

int gm_scalars_upload(gm_scalars** out, const uint64_t* host, int n, int limbs, uint32_t gpu_index);
void gm_scalars_free(gm_scalars* s);

/* MSM over resident data: bases[base_offset .. +n) with scalars
 * [scalar_offset .. +n). Output: affine ABI record (G1Point / G2Point),
 * Montgomery form. scalar_bits: bound on scalar values (0 = 64 * limbs).
 * window <= 0 picks the heuristic. device_combine != 0 runs the final
 * window combine on the GPU instead of the host. */
int gm_msm_resident(void* result_affine, const gm_bases* b, int base_offset, const gm_scalars* s, int scalar_offset,
                    int n, int scalar_bits, int window, int device_combine);

/* One MSM of a batch over resident data. window/seg <= 0: defaults. */
typedef struct gm_msm_spec {
    int n;
    int base_offset;
    int scalar_offset;
    int window; /* window bits c */
    int seg;    /* buckets per bucket-reduction lane (any count >= 1; capped at 2^(window-1)) */
} gm_msm_spec;

/* Many MSMs of uneven sizes in one pipeline run: every stage is one launch
 * over all MSMs (one sort, one balanced accumulation over all entries, one
 * reduction), then one window combine per MSM. results_affine: `count` ABI
 * affine records. */
int gm_msm_batch(void* results_affine, const gm_bases* b, const gm_scalars* s, const gm_msm_spec* specs, int count,
                 int scalar_bits, int device_combine);

/* A fixed batch bound to resident bases and scalars, with its own stream and
 * device workspace. GM_BATCH_GRAPH captures the device pipeline (sort,
 * accumulation, reduction, window sums and their copy to pinned host memory)
 * in a hipGraph at creation; every launch replays it. gm_batch_wait then runs
 * the window combines on host threads, or, with GM_BATCH_DEVICE_COMBINE, on
 * the device inside the same graph (one lane per MSM). The scalars' contents
 * may change between runs; the handles must outlive the batch. One batch must
 * not be launched again before gm_batch_wait returns; different batches may
 * overlap on the device. */
#define GM_BATCH_GRAPH 0x1
#define GM_BATCH_DEVICE_COMBINE 0x2
typedef struct gm_batch gm_batch;
int gm_batch_create(gm_batch** out, const gm_bases* b, const gm_scalars* s, const gm_msm_spec* specs, int count,
                    int scalar_bits, int flags);
int gm_batch_launch(gm_batch* bt);                       /* asynchronous */
int gm_batch_wait(gm_batch* bt, void* results_affine);   /* count ABI affine records */
int gm_batch_run(gm_batch* bt, void* results_affine);    /* launch + wait */
int gm_batch_count(const gm_batch* bt);
size_t gm_batch_workspace_bytes(const gm_batch* bt);
void gm_batch_free(gm_batch* bt);

/* ------------------------------------------------------------ planner MLP */

/* FP32 MLP 3 -> 32 -> 32 (ReLU) -> 2 used by the MSM planner; parameters are
 * W1[32][3] b1[32] W2[32][32] b2[32] W3[2][32] b3[2]. x: n x 3, y/t: n x 2.
 * Loss L = (1/n) sum 0.5 (y - t)^2. */
#define GM_MLP_PARAMS 1250
int gm_mlp_forward(const float* params, const float* x, int n, float* y, uint32_t gpu_index);
/* Full-batch gradient dL/dparams and L at `params`. */
int gm_mlp_grad(const float* params, const float* x, const float* t, int n, float* grad, float* loss,
                uint32_t gpu_index);
/* `epochs` full-batch Adam steps (one step captured in a hipGraph and
 * replayed); params updated in place; loss_curve (optional, epochs floats)
 * receives the loss before each step. */
int gm_mlp_train(float* params, const float* x, const float* t, int n, int epochs, float lr, float* loss_curve,
                 uint32_t gpu_index);

/* --------------------------------------------------------- request server */

/* Continuous batching of mixed G1/G2 MSM requests over resident base pools.
 * A scheduler thread forms batches from the head of the queue (one curve per
 * batch) whenever one of `slots` stream slots is free, admitting requests
 * while the batch's device footprint fits the HBM budget left by in-flight
 * batches; batches on different slots overlap on the device; a worker per
 * slot stages each batch's scalars and frees the slot when the device is done,
 * and a pool of host threads finishes the window combines. */
#define GM_PLANNER_HEURISTIC 0
#define GM_PLANNER_MLP 1
typedef struct gm_server_config {
    uint32_t gpu_index;
    int slots;                /* batches in flight at once (1..16) */
    int max_batch;            /* MSMs per batch */
    size_t hbm_budget_bytes;  /* admission limit for in-flight footprints */
    int planner;              /* GM_PLANNER_* */
    const float* mlp_params;  /* GM_MLP_PARAMS floats when planner == MLP */
    int small_max_n;          /* requests of at most this many points form their own batches, go first and keep one
                                 slot to themselves; 0: 2^16, < 0: no small class */
} gm_server_config;

typedef struct gm_server_stats {
    uint64_t submitted, completed, failed, batches;
    double mean_batch;           /* MSMs per batch */
    double p50_ms, p99_ms, max_ms; /* submit -> result latency of completed requests */
    size_t peak_inflight_bytes;
    int peak_concurrent_batches;
} gm_server_stats;

typedef struct gm_server gm_server;
int gm_server_create(gm_server** out, const gm_server_config* cfg, const gm_bases* g1_pool, const gm_bases* g2_pool);
/* Queues sum_i scalars[i] * pool[base_offset + i]. The scalars are copied,
 * or with GM_SUBMIT_BORROW read in place (the caller keeps them unchanged
 * until gm_server_wait returns for this ticket). */
#define GM_SUBMIT_BORROW 0x1
int gm_server_submit(gm_server* s, int curve, int base_offset, int n, const uint64_t* scalars, int limbs,
                     int scalar_bits, uint32_t flags, uint64_t* ticket);
/* Waits for a ticket (timeout_ms < 0: forever); 0 and the affine ABI record
 * on success, the request's error code otherwise, -6 on timeout. */
int gm_server_wait(gm_server* s, uint64_t ticket, void* result_affine, double* latency_ms, int timeout_ms);
int gm_server_get_stats(gm_server* s, gm_server_stats* out);
/* Test hook: the next `batches` batches fail as if the device had failed
 * (no kernel is launched); their requests complete with -4. */
int gm_server_inject_failures(gm_server* s, int batches);
/* Drains queued and in-flight requests, then stops. */
void gm_server_destroy(gm_server* s);

/* ------------------------------------------------- MSM without allocation */

/* No counterpart in the reference. For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the MSM of point_msm_async<PointType> (curve.h) over device
 * pointers, d_points n affine ABI records (G1Point / G2Point) in Montgomery
 * form, d_scalars n x scalar_limbs little-endian u64 words, d_result one
 * canonical affine ABI record, all enqueued on `stream` (a hipStream_t; NULL
 * for the null stream) with no host wait. It calls no hipMalloc / hipFree or
 * allocating library routine: the workspace is one fixed 144 MiB device array
 * per GPU, held until the process ends, and the window combine runs on the
 * device. Sizes whose workspace does not fit (about 2^16 G1 or 2^15 G2 points
 * with 5-limb scalars) return -6. Calls that overlap on one GPU share that
 * array and race: serialise them. Returns 0; -1 bad arguments; -2 device or
 * set-up failure; -4 launch failure; -6 too large for the fixed workspace. */
int gm_point_msm_noalloc(void* stream, uint32_t gpu_index, int curve, void* d_result, const void* d_points,
                         const uint64_t* d_scalars, int scalar_limbs, int n);

/* ------------------------------------------------------- host closed form */

/* k * G on the host for the curve's generator G (G1 or G2), k given as
 * `limbs` little-endian u64 words: the affine ABI record (G1Point / G2Point,
 * Montgomery form, canonical; O as (0, 0, infinity)). A workload whose bases
 * are P_i = k_i G has MSM result (sum_i s_i k_i mod r) G, so this checks a
 * GPU MSM against an independent host computation. Returns 0, or -1 for a
 * bad curve / limbs outside [1, 8]. */
int gm_generator_multiple(int curve, const uint64_t* k, int limbs, void* result_affine);

/* -------------------------------------------------------------- profiling */

/* Record HIP events between pipeline stages of subsequent MSMs on this GPU. */
int gm_set_profiling(uint32_t gpu_index, int on);
/* Plan of the last MSM run on this GPU: window bits, windows, and the number
 * of (window, bucket) entries the accumulation kernel added. */
int gm_last_msm_info(uint32_t gpu_index, int* window, int* windows, long long* entries);
/* Stage times (ms) of the last profiled MSM: [hist, scan, scatter,
 * accumulate, fixup, bucket_reduce, window_sum]. */
int gm_profile_last(uint32_t gpu_index, float* stage_ms, int count);

/* ---------------------------------------------------------- point formats
 * The wire format on the caller's side of the MSM (reference
 * rust-api/src/conversions.rs:57-79 converts between the ABI records and
 * tfhe-zk-pok's arkworks points). All pointers are HOST pointers; the work
 * runs on the GPU in one batch. Points are ABI affine records in Montgomery
 * form with field values below p, as every MSM entry point takes and returns.
 *
 * Encoding (the layout of ark-serialize's CanonicalSerialize for
 * short-Weierstrass affine points): each Fp is its canonical integer in 56
 * little-endian bytes, an Fp2 is c0 then c1; compressed = x, uncompressed =
 * x then y; the last byte of the encoding carries the flags: 0x40 point at
 * infinity (coordinates zero), 0x80 y is the larger of {y, -y} (integers for
 * Fp; for Fp2, c1 decides and c0 breaks a tie). Sizes: G1 56 / 112 bytes,
 * G2 112 / 224 bytes (compressed / uncompressed).
 *
 * Status per point (deserialize, check): 0 ok, 1 invalid encoding (a field
 * element >= p, or both flag bits), 2 not on the curve (compressed: x^3 + b
 * is not a square), 3 not in the prime-order subgroup. Points that fail are
 * written as the point at infinity. The functions return the number of points
 * with a non-zero status (0 = all valid), or -1 bad arguments, -2 no device,
 * -3 allocation failure, -4 HIP error. No counterpart in the reference. */
#define GM_CHECK_ON_CURVE 0x1 /* y^2 = x^3 + b with b = 1 (G1), 1 + i (G2) */
#define GM_CHECK_SUBGROUP 0x2 /* r * P = O */

int gm_point_encoded_size(int curve, int compressed);
int gm_points_serialize(int curve, void* out_bytes, const void* points, int n, int compressed, uint32_t gpu_index);
/* Compressed input is decompressed with a square root (then on the curve by
 * construction); `checks` selects the validation of uncompressed input and
 * the subgroup check of either. */
int gm_points_deserialize(int curve, void* points, const void* in_bytes, int n, int compressed, int checks,
                          int32_t* status, uint32_t gpu_index);
/* Validation of ABI records (the reference has no is_on_curve; its curve
 * constants name b = 4 and 4i, reference src/curve.cu:280-289, while the
 * generators lie on y^2 = x^3 + 1 and y^2 = x^3 + (1 + i)). */
int gm_points_check(int curve, const void* points, int n, int checks, int32_t* status, uint32_t gpu_index);
/* Montgomery records <-> canonical standard-form records (the limbs
 * from_sign_and_limbs takes and into_bigint returns); infinity records are
 * written as x = y = 0 with the flag set. */
int gm_points_to_canonical(int curve, void* out, const void* points, int n, uint32_t gpu_index);
int gm_points_from_canonical(int curve, void* out, const void* points, int n, uint32_t gpu_index);

/* ------------------------------------------------------------ diagnostics */

/* No counterpart in the reference; for tests of the bucket kernels' lazily
 * reduced field products. Runs one product per element on device arrays of n
 * elements of 16 uint32 28-bit limbs (G1, the Fp products, one element per
 * lane) or 2 x 16 (G2, c0 then c1: the Fp2 products on the lane pairs those
 * kernels use, one element per pair), Montgomery form; chain != 0 selects the
 * accumulation's single-chain column sums, else the fix-up and tail kernels'
 * form. op 0 a b, 1 a b (a, b raw), 2 a b - c d, 3 a b - c d (a, b raw),
 * 4 a^2 (a raw), 5 a^2 - b - 2c (a raw). Operands lie in [0, 2p); raw ones
 * are limb-wise differences of two such values. Results lie in [0, 2p) with
 * normalised limbs. Asynchronous on `stream` (a hipStream_t); returns a
 * hipError_t code. */
int gm_field_ops_check(int curve, int chain, int op, const void* a, const void* b, const void* c, const void* d,
                       void* out, int n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GPU_MSM_EXT_H */
