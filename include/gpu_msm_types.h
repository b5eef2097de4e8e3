/*
 * ABI structs of the BLS12-446 MSM library (C and C++).
 *
 * Layouts are byte-identical to the reference's (reference include/fp.h:13-15,
 * include/fp2.h:10-13, include/curve.h:13-43, mirrored by
 * rust-api/src/ffi.rs:8-57): little-endian 64-bit limbs, Montgomery form with
 * R = 2^448 unless a function says otherwise, `bool infinity` padded to 8 bytes.
 */
#ifndef GPU_MSM_TYPES_H
#define GPU_MSM_TYPES_H

#include <stdbool.h>
#include <stdint.h>

#define FP_LIMBS 7
#define FP_BITS 446

typedef struct Fp {
    uint64_t limb[FP_LIMBS];
} Fp;

typedef struct Fp2 {
    Fp c0; /* real part */
    Fp c1; /* coefficient of i, i^2 = -1 */
} Fp2;

typedef struct G1Point {
    Fp x;
    Fp y;
    bool infinity;
} G1Point;

typedef struct G2Point {
    Fp2 x;
    Fp2 y;
    bool infinity;
} G2Point;

/* Standard projective (X:Y:Z), x = X/Z, y = Y/Z, Z = 0 is the point at infinity. */
typedef struct G1ProjectivePoint {
    Fp X;
    Fp Y;
    Fp Z;
} G1ProjectivePoint;

typedef struct G2ProjectivePoint {
    Fp2 X;
    Fp2 Y;
    Fp2 Z;
} G2ProjectivePoint;

#ifdef __cplusplus
static_assert(sizeof(Fp) == 56, "Fp layout");
static_assert(sizeof(Fp2) == 112, "Fp2 layout");
static_assert(sizeof(G1Point) == 120, "G1Point layout");
static_assert(sizeof(G2Point) == 232, "G2Point layout");
static_assert(sizeof(G1ProjectivePoint) == 168, "G1ProjectivePoint layout");
static_assert(sizeof(G2ProjectivePoint) == 336, "G2ProjectivePoint layout");
#endif

#endif /* GPU_MSM_TYPES_H */
