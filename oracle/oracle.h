/*
 * CPU oracle for the BLS12-446 G1/G2 MSM path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * host arithmetic, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. The product library never
 * links, loads or calls it.
 *
 * Struct layouts are the reference's ABI layouts (include/curve.h:13-43 of the
 * reference): Fp = 7 x u64 little-endian (56 B), Fp2 = {c0, c1} (112 B),
 * G1Point = {x, y, bool infinity} (120 B), G2Point (232 B),
 * G1ProjectivePoint = {X, Y, Z} (168 B), G2ProjectivePoint (336 B).
 * Field elements are canonical (< p) and, unless a name says otherwise, in
 * Montgomery form with R = 2^448.
 */
#ifndef GPU_MSM_ORACLE_H
#define GPU_MSM_ORACLE_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t limb[7]; } OFp;
typedef struct { OFp c0, c1; } OFp2;
typedef struct { OFp x, y; bool infinity; } OG1;
typedef struct { OFp2 x, y; bool infinity; } OG2;
typedef struct { OFp X, Y, Z; } OG1P;
typedef struct { OFp2 X, Y, Z; } OG2P;

/* Fp (reference src/fp.cu) */
int  or_fp_cmp(const OFp* a, const OFp* b);
void or_fp_add(OFp* c, const OFp* a, const OFp* b);
void or_fp_sub(OFp* c, const OFp* a, const OFp* b);
void or_fp_neg(OFp* c, const OFp* a);
void or_fp_mont_mul(OFp* c, const OFp* a, const OFp* b);
void or_fp_to_mont(OFp* c, const OFp* a);
void or_fp_from_mont(OFp* c, const OFp* a);
void or_fp_mont_inv(OFp* c, const OFp* a);
void or_fp_modulus(OFp* p);
void or_fp_one_mont(OFp* one);

/* Fp2 = Fp[i]/(i^2+1) (reference src/fp2.cu) */
void or_fp2_add(OFp2* c, const OFp2* a, const OFp2* b);
void or_fp2_sub(OFp2* c, const OFp2* a, const OFp2* b);
void or_fp2_mont_mul(OFp2* c, const OFp2* a, const OFp2* b);
void or_fp2_mont_inv(OFp2* c, const OFp2* a);

/* Curve (reference src/curve.cu) */
void or_g1_generator(OG1* g_standard);
void or_g2_generator(OG2* g_standard);
void or_g1_to_mont(OG1* out, const OG1* in);
void or_g2_to_mont(OG2* out, const OG2* in);
void or_g1_from_mont(OG1* out, const OG1* in);
void or_g2_from_mont(OG2* out, const OG2* in);
void or_g1_affine_to_proj(OG1P* out, const OG1* in);
void or_g2_affine_to_proj(OG2P* out, const OG2* in);
void or_g1_proj_to_affine(OG1* out, const OG1P* in);
void or_g2_proj_to_affine(OG2* out, const OG2P* in);
void or_g1_proj_add(OG1P* r, const OG1P* p1, const OG1P* p2);
void or_g2_proj_add(OG2P* r, const OG2P* p1, const OG2P* p2);
void or_g1_proj_double(OG1P* r, const OG1P* p);
void or_g2_proj_double(OG2P* r, const OG2P* p);
bool or_g1_is_on_curve(const OG1* p_mont);   /* y^2 = x^3 + 1 */
bool or_g2_is_on_curve(const OG2* p_mont);   /* y^2 = x^3 + (1 + i) */

/* k * P by projective double-and-add, k little-endian u64 limbs; affine out. */
void or_g1_scalar_mul(OG1* out, const OG1* base, const uint64_t* k, int limbs);
void or_g2_scalar_mul(OG2* out, const OG2* base, const uint64_t* k, int limbs);

/* Bases P_i = k_i * G (Montgomery affine), one 64-bit k_i per base. */
void or_g1_bases_from_k(OG1* out, const uint64_t* k, int n, int threads);
void or_g2_bases_from_k(OG2* out, const uint64_t* k, int n, int threads);

/* Pippenger MSM on the host: sum_i s_i * P_i, scalars at stride `limbs` u64.
 * c = window bits (0 picks a default), threads = OpenMP threads (1 = serial).
 * Output: affine, Montgomery form. */
void or_g1_msm(OG1* out, const OG1* bases, const uint64_t* scalars, int limbs, int n, int c, int threads);
void or_g2_msm(OG2* out, const OG2* bases, const uint64_t* scalars, int limbs, int n, int c, int threads);

/* std::mt19937_64 restated (published Matsumoto-Nishimura algorithm). */
typedef struct { uint64_t mt[312]; int idx; } OMt64;
void     or_mt64_seed(OMt64* s, uint64_t seed);
uint64_t or_mt64_next(OMt64* s);
/* n uniform 64-bit draws. */
void or_gen_u64(uint64_t seed, int n, uint64_t* out);
/* n scalars uniform in [0, r): 5 draws per candidate, top limb masked to
 * 43 bits, rejected when >= r. Output 5 limbs per scalar. */
void or_gen_scalars_modr(uint64_t seed, int n, uint64_t* out);
void or_scalar_order(uint64_t r_out[5]);

#ifdef __cplusplus
}
#endif
#endif
