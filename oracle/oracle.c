/*
 * CPU oracle for the BLS12-446 G1/G2 MSM path — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference host arithmetic. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this; the
 * product library never does.
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   Fp   — Python big-int on random vectors, the reference test_fp.cu known
 *          answers (1+1, 5*3, (p-1)+1, 50-100, (p-1)^2), Montgomery round trip.
 *   Fp2  — Python big-int over Fp[i]/(i^2+1); (1+i)^2 = 2i, i^2 = -1.
 *   curve— textbook affine group law in Python, on-curve (b=1, b'=1+i),
 *          r*G = O, projective vs affine.
 *   MSM  — closed form (sum s_i k_i mod r)*G, triangular-number closed form of
 *          the reference test_msm.cu, brute force on small inputs.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* p, reference src/fp.cu:15-19 (DEVICE_MODULUS). */
static const OFp P = {{0x311c0026aab0aaabULL, 0x56ee4528c573b5ccULL, 0x824e6dc3e23acdeeULL,
                       0x0f75a64bbac71602ULL, 0x0095a4b78a02fe32ULL, 0x200fc34965aad640ULL,
                       0x3cdee0fb28c5e535ULL}};
/* R^2 mod p, reference src/fp.cu:21-25 (DEVICE_R2). */
static const OFp R2 = {{0x2AFF01DDDC752B45ULL, 0x92C772A7421CCF5BULL, 0x140EEF29C347DAD6ULL,
                        0xF5A1400C22EA595EULL, 0x99D91C9FEC145218ULL, 0x3BB6537F90143D4BULL,
                        0x3627854C9BE7974FULL}};
/* -p^{-1} mod 2^64, reference src/fp.cu:33 (DEVICE_P_PRIME). */
static const uint64_t PPRIME = 0xcd63fd900035fffdULL;
/* Scalar-field order r = x^4 - x^2 + 1 for the BLS12 seed x = -0x6008204000000020001. */
static const uint64_t RORD[5] = {0x0428001400040001ULL, 0x7bb9b0e8d8ca3461ULL, 0xd04c98ccc4c050bcULL,
                                 0x7995b34995830fa4ULL, 0x00000511b70539f2ULL};

/* ------------------------------------------------------------------------- Fp */

/* reference src/fp.cu:115-122 */
int or_fp_cmp(const OFp* a, const OFp* b) {
    for (int i = 6; i >= 0; i--) {
        if (a->limb[i] > b->limb[i]) return 1;
        if (a->limb[i] < b->limb[i]) return -1;
    }
    return 0;
}

static int fp_is_zero(const OFp* a) {
    uint64_t o = 0;
    for (int i = 0; i < 7; i++) o |= a->limb[i];
    return o == 0;
}

static uint64_t add_raw(OFp* c, const OFp* a, const OFp* b) {
    u128 carry = 0;
    for (int i = 0; i < 7; i++) {
        carry += (u128)a->limb[i] + b->limb[i];
        c->limb[i] = (uint64_t)carry;
        carry >>= 64;
    }
    return (uint64_t)carry;
}

static uint64_t sub_raw(OFp* c, const OFp* a, const OFp* b) {
    uint64_t borrow = 0;
    for (int i = 0; i < 7; i++) {
        u128 d = (u128)a->limb[i] - b->limb[i] - borrow;
        c->limb[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
    return borrow;
}

/* reference src/fp.cu:181-194: one conditional subtraction of p. */
void or_fp_add(OFp* c, const OFp* a, const OFp* b) {
    OFp s;
    uint64_t carry = add_raw(&s, a, b);
    if (carry || or_fp_cmp(&s, &P) >= 0) sub_raw(&s, &s, &P);
    *c = s;
}

/* reference src/fp.cu:197-208: add p back on borrow. */
void or_fp_sub(OFp* c, const OFp* a, const OFp* b) {
    OFp d;
    if (sub_raw(&d, a, b)) add_raw(&d, &d, &P);
    *c = d;
}

/* reference src/fp.cu:436-443 */
void or_fp_neg(OFp* c, const OFp* a) {
    if (fp_is_zero(a)) {
        memset(c, 0, sizeof *c);
        return;
    }
    sub_raw(c, &P, a);
}

/* Montgomery product a*b*2^-448 mod p. Schoolbook 7x7 product followed by a
 * word-by-word Montgomery reduction and one conditional subtraction, as in
 * reference src/fp.cu:249-292 (fp_mul_raw) and :306-366 (fp_mont_reduce). */
static void mont_reduce(OFp* c, uint64_t t[15]) {
    for (int i = 0; i < 7; i++) {
        uint64_t u = t[i] * PPRIME;
        u128 carry = 0;
        for (int j = 0; j < 7; j++) {
            carry += (u128)u * P.limb[j] + t[i + j];
            t[i + j] = (uint64_t)carry;
            carry >>= 64;
        }
        for (int k = i + 7; carry && k < 15; k++) {
            carry += t[k];
            t[k] = (uint64_t)carry;
            carry >>= 64;
        }
    }
    OFp r;
    memcpy(r.limb, t + 7, sizeof r.limb);
    if (t[14] || or_fp_cmp(&r, &P) >= 0) sub_raw(&r, &r, &P);
    *c = r;
}

void or_fp_mont_mul(OFp* c, const OFp* a, const OFp* b) {
    uint64_t t[15] = {0};
    for (int i = 0; i < 7; i++) {
        u128 carry = 0;
        for (int j = 0; j < 7; j++) {
            carry += (u128)a->limb[i] * b->limb[j] + t[i + j];
            t[i + j] = (uint64_t)carry;
            carry >>= 64;
        }
        t[i + 7] = (uint64_t)carry;
    }
    mont_reduce(c, t);
}

/* reference src/fp.cu:378-386: a*R^2*R^-1 */
void or_fp_to_mont(OFp* c, const OFp* a) { or_fp_mont_mul(c, a, &R2); }

/* reference src/fp.cu:390-405: reduce [a, 0] */
void or_fp_from_mont(OFp* c, const OFp* a) {
    uint64_t t[15] = {0};
    memcpy(t, a->limb, sizeof a->limb);
    mont_reduce(c, t);
}

void or_fp_modulus(OFp* p) { *p = P; }

/* reference src/fp.cu:132-139: R mod p */
void or_fp_one_mont(OFp* one) {
    OFp o = {{1, 0, 0, 0, 0, 0, 0}};
    or_fp_to_mont(one, &o);
}

/* reference src/fp.cu:545-568: Fermat a^(p-2); returns 0 for 0. Any exponent
 * chain gives the same canonical value, so the chain here is plain binary. */
void or_fp_mont_inv(OFp* c, const OFp* a) {
    if (fp_is_zero(a)) {
        memset(c, 0, sizeof *c);
        return;
    }
    OFp e = P;
    e.limb[0] -= 2; /* p is odd and > 2; limb 0 of p is >= 2 */
    OFp acc;
    or_fp_one_mont(&acc);
    for (int i = 6; i >= 0; i--) {
        for (int b = 63; b >= 0; b--) {
            or_fp_mont_mul(&acc, &acc, &acc);
            if ((e.limb[i] >> b) & 1) or_fp_mont_mul(&acc, &acc, a);
        }
    }
    *c = acc;
}

/* ------------------------------------------------------------------------ Fp2 */

/* reference src/fp2.cu:40-50 */
void or_fp2_add(OFp2* c, const OFp2* a, const OFp2* b) {
    or_fp_add(&c->c0, &a->c0, &b->c0);
    or_fp_add(&c->c1, &a->c1, &b->c1);
}
void or_fp2_sub(OFp2* c, const OFp2* a, const OFp2* b) {
    or_fp_sub(&c->c0, &a->c0, &b->c0);
    or_fp_sub(&c->c1, &a->c1, &b->c1);
}

/* reference src/fp2.cu:88-112: Karatsuba with i^2 = -1.
 * t0 = a0 b0, t1 = a1 b1, t2 = (a0+a1)(b0+b1); c0 = t0 - t1, c1 = t2 - t0 - t1 */
void or_fp2_mont_mul(OFp2* c, const OFp2* a, const OFp2* b) {
    OFp t0, t1, t2, t3, r0, r1;
    or_fp_mont_mul(&t0, &a->c0, &b->c0);
    or_fp_mont_mul(&t1, &a->c1, &b->c1);
    or_fp_add(&t2, &a->c0, &a->c1);
    or_fp_add(&t3, &b->c0, &b->c1);
    or_fp_mont_mul(&t2, &t2, &t3);
    or_fp_sub(&r0, &t0, &t1);
    or_fp_sub(&r1, &t2, &t0);
    or_fp_sub(&r1, &r1, &t1);
    c->c0 = r0;
    c->c1 = r1;
}

static int fp2_is_zero(const OFp2* a) { return fp_is_zero(&a->c0) && fp_is_zero(&a->c1); }

/* reference src/fp2.cu:232-259: conj(a) / (a0^2 + a1^2) */
void or_fp2_mont_inv(OFp2* c, const OFp2* a) {
    if (fp2_is_zero(a)) {
        memset(c, 0, sizeof *c);
        return;
    }
    OFp t0, t1, n, ni, r0, r1;
    or_fp_mont_mul(&t0, &a->c0, &a->c0);
    or_fp_mont_mul(&t1, &a->c1, &a->c1);
    or_fp_add(&n, &t0, &t1);
    or_fp_mont_inv(&ni, &n);
    or_fp_mont_mul(&r0, &a->c0, &ni);
    or_fp_neg(&r1, &a->c1);
    or_fp_mont_mul(&r1, &r1, &ni);
    c->c0 = r0;
    c->c1 = r1;
}

/* ---------------------------------------------------------------------- curve */

/* Generators in standard form, reference src/curve.cu:786-814. */
void or_g1_generator(OG1* g) {
    static const OG1 G = {
        {{0x3bf9166c8236f4faULL, 0x8bc02b7cbe6a9e8dULL, 0x11c1e56b3e4bc80bULL, 0x6b20d782901a6f62ULL,
          0x2ce8c34265bf3841ULL, 0x11b73d3d76ae9851ULL, 0x326ed6bd777fc6a3ULL}},
        {{0xfe6f792612016b30ULL, 0x22db0ce6034a9db9ULL, 0xb9093f32002756daULL, 0x39d7f424b6660204ULL,
          0xf843c947aa57f571ULL, 0xd6d62d244e413636ULL, 0x1a7caf4a4d3887a6ULL}},
        false};
    *g = G;
}

void or_g2_generator(OG2* g) {
    static const OG2 G = {
        {{{0x0e529ee4dce9991dULL, 0xd6ebaf149094f1ccULL, 0x043c6bf16312d638ULL, 0x062b61439640e885ULL,
           0x18dad8ed784dd225ULL, 0xa57c0038441f7d15ULL, 0x21f8d4a76f74541aULL}},
         {{0xcaf5185423a7d23aULL, 0x7cef6acb145b6413ULL, 0x2879dd439b019b8bULL, 0x71449cdeca4f0007ULL,
           0xdebaf4a2c5534527ULL, 0xa1b4e791d1b86560ULL, 0x1e0f563c601bb8dcULL}}},
        {{{0x274315837455b919ULL, 0x82039e4221ff3507ULL, 0x00346cebad16a036ULL, 0x0177bfd6654e681eULL,
           0xddff621b5db3f897ULL, 0x0cc61570301497a7ULL, 0x115ea2305a78f646ULL}},
         {{0x392236e9cf2976c2ULL, 0xd8ab17c84b9f03cdULL, 0x8a8e6755f9d82fd1ULL, 0x7532834528cd5a64ULL,
           0x0b0bcc3fb6f2161cULL, 0x76a2ffcb7d47679dULL, 0x25ed2192b203c1feULL}}},
        false};
    *g = G;
}

void or_g1_to_mont(OG1* out, const OG1* in) {
    OG1 r = *in;
    if (!in->infinity) {
        or_fp_to_mont(&r.x, &in->x);
        or_fp_to_mont(&r.y, &in->y);
    }
    *out = r;
}
void or_g2_to_mont(OG2* out, const OG2* in) {
    OG2 r = *in;
    if (!in->infinity) {
        or_fp_to_mont(&r.x.c0, &in->x.c0);
        or_fp_to_mont(&r.x.c1, &in->x.c1);
        or_fp_to_mont(&r.y.c0, &in->y.c0);
        or_fp_to_mont(&r.y.c1, &in->y.c1);
    }
    *out = r;
}
/* reference rust-api/src/c_wrapper.cu:261-283: infinity maps to (0, 0, true). */
void or_g1_from_mont(OG1* out, const OG1* in) {
    OG1 r;
    memset(&r, 0, sizeof r);
    if (in->infinity) {
        r.infinity = true;
    } else {
        or_fp_from_mont(&r.x, &in->x);
        or_fp_from_mont(&r.y, &in->y);
    }
    *out = r;
}
void or_g2_from_mont(OG2* out, const OG2* in) {
    OG2 r;
    memset(&r, 0, sizeof r);
    if (in->infinity) {
        r.infinity = true;
    } else {
        or_fp_from_mont(&r.x.c0, &in->x.c0);
        or_fp_from_mont(&r.x.c1, &in->x.c1);
        or_fp_from_mont(&r.y.c0, &in->y.c0);
        or_fp_from_mont(&r.y.c1, &in->y.c1);
    }
    *out = r;
}

/* reference src/curve.cu:396-432: infinity -> (0,0,0), else Z = R mod p. */
void or_g1_affine_to_proj(OG1P* out, const OG1* in) {
    memset(out, 0, sizeof *out);
    if (in->infinity) return;
    out->X = in->x;
    out->Y = in->y;
    or_fp_one_mont(&out->Z);
}
void or_g2_affine_to_proj(OG2P* out, const OG2* in) {
    memset(out, 0, sizeof *out);
    if (in->infinity) return;
    out->X = in->x;
    out->Y = in->y;
    or_fp_one_mont(&out->Z.c0);
}

/* reference src/curve.cu:754-778: Z = 0 -> infinity (0, 0, true). */
void or_g1_proj_to_affine(OG1* out, const OG1P* in) {
    OG1 r;
    memset(&r, 0, sizeof r);
    if (fp_is_zero(&in->Z)) {
        r.infinity = true;
    } else {
        OFp zi;
        or_fp_mont_inv(&zi, &in->Z);
        or_fp_mont_mul(&r.x, &in->X, &zi);
        or_fp_mont_mul(&r.y, &in->Y, &zi);
    }
    *out = r;
}
void or_g2_proj_to_affine(OG2* out, const OG2P* in) {
    OG2 r;
    memset(&r, 0, sizeof r);
    if (fp2_is_zero(&in->Z)) {
        r.infinity = true;
    } else {
        OFp2 zi;
        or_fp2_mont_inv(&zi, &in->Z);
        or_fp2_mont_mul(&r.x, &in->X, &zi);
        or_fp2_mont_mul(&r.y, &in->Y, &zi);
    }
    *out = r;
}

/* The projective formulas below are written once over a tiny "field vtable"
 * so G1 and G2 share one statement of reference src/curve.cu:444-738. */
typedef void (*fbin)(void*, const void*, const void*);
typedef struct {
    size_t sz;
    fbin add, sub, mul;
    int (*is_zero)(const void*);
    void (*small)(void*, int); /* Montgomery form of a small integer */
} Field;

static void fp_small(void* o, int v) {
    OFp t = {{(uint64_t)v, 0, 0, 0, 0, 0, 0}};
    or_fp_to_mont((OFp*)o, &t);
}
static void fp2_small(void* o, int v) {
    OFp2* z = (OFp2*)o;
    memset(z, 0, sizeof *z);
    fp_small(&z->c0, v);
}
static int fp_is_zero_v(const void* a) { return fp_is_zero((const OFp*)a); }
static int fp2_is_zero_v(const void* a) { return fp2_is_zero((const OFp2*)a); }

static const Field F1 = {sizeof(OFp), (fbin)or_fp_add, (fbin)or_fp_sub, (fbin)or_fp_mont_mul, fp_is_zero_v, fp_small};
static const Field F2 = {sizeof(OFp2), (fbin)or_fp2_add, (fbin)or_fp2_sub, (fbin)or_fp2_mont_mul, fp2_is_zero_v,
                         fp2_small};

/* Element i of a projective point stored as {X, Y, Z}. */
#define PX(pt) ((unsigned char*)(pt))
#define PY(pt, f) ((unsigned char*)(pt) + (f)->sz)
#define PZ(pt, f) ((unsigned char*)(pt) + 2 * (f)->sz)

typedef struct { uint64_t w[14]; } Tmp; /* large enough for an Fp2 */

/* reference src/curve.cu:591-657 (G1) / 660-738 (G2): a = 0 doubling
 * A = 3X^2, B = YZ, C = XYB, D = A^2 - 8C, X3 = 2BD,
 * Y3 = A(4C - D) - 8Y^2B^2, Z3 = 8B^3. */
static void proj_double(const Field* f, void* r, const void* p) {
    if (f->is_zero(PZ(p, f))) {
        memmove(r, p, 3 * f->sz);
        return;
    }
    Tmp c2, c3, c4, c8, xsq, A, B, XY, C, Asq, C8, D, BD, C4, t, Ysq, Bsq, YB, Bc, X3, Y3, Z3;
    f->small(&c2, 2);
    f->small(&c3, 3);
    f->small(&c4, 4);
    f->small(&c8, 8);
    f->mul(&xsq, PX(p), PX(p));
    f->mul(&A, &c3, &xsq);
    f->mul(&B, PY(p, f), PZ(p, f));
    f->mul(&XY, PX(p), PY(p, f));
    f->mul(&C, &XY, &B);
    f->mul(&Asq, &A, &A);
    f->mul(&C8, &c8, &C);
    f->sub(&D, &Asq, &C8);
    f->mul(&BD, &B, &D);
    f->mul(&X3, &c2, &BD);
    f->mul(&C4, &c4, &C);
    f->sub(&t, &C4, &D);
    f->mul(&t, &A, &t);
    f->mul(&Ysq, PY(p, f), PY(p, f));
    f->mul(&Bsq, &B, &B);
    f->mul(&YB, &Ysq, &Bsq);
    f->mul(&YB, &c8, &YB);
    f->sub(&Y3, &t, &YB);
    f->mul(&Bc, &Bsq, &B);
    f->mul(&Z3, &c8, &Bc);
    memcpy(PX(r), &X3, f->sz);
    memcpy(PY(r, f), &Y3, f->sz);
    memcpy(PZ(r, f), &Z3, f->sz);
}

/* reference src/curve.cu:444-519 (G1) / 522-588 (G2):
 * u = Y2Z1 - Y1Z2, v = X2Z1 - X1Z2; u = v = 0 -> double;
 * R = v^2 X1Z2, A = u^2 Z1Z2 - v^3 - 2R,
 * X3 = vA, Y3 = u(R - A) - v^3 Y1Z2, Z3 = v^3 Z1Z2. */
static void proj_add(const Field* f, void* r, const void* p1, const void* p2) {
    if (f->is_zero(PZ(p1, f))) {
        memmove(r, p2, 3 * f->sz);
        return;
    }
    if (f->is_zero(PZ(p2, f))) {
        memmove(r, p1, 3 * f->sz);
        return;
    }
    Tmp Y1Z2, X1Z2, Z1Z2, Y2Z1, X2Z1, u, uu, v, vv, vvv, R, A, t1, t2, c2, R2x, RA, X3, Y3, Z3;
    f->mul(&Y1Z2, PY(p1, f), PZ(p2, f));
    f->mul(&X1Z2, PX(p1), PZ(p2, f));
    f->mul(&Z1Z2, PZ(p1, f), PZ(p2, f));
    f->mul(&Y2Z1, PY(p2, f), PZ(p1, f));
    f->sub(&u, &Y2Z1, &Y1Z2);
    f->mul(&uu, &u, &u);
    f->mul(&X2Z1, PX(p2), PZ(p1, f));
    f->sub(&v, &X2Z1, &X1Z2);
    if (f->is_zero(&u) && f->is_zero(&v)) {
        proj_double(f, r, p1);
        return;
    }
    f->mul(&vv, &v, &v);
    f->mul(&vvv, &v, &vv);
    f->mul(&R, &vv, &X1Z2);
    f->mul(&t1, &uu, &Z1Z2);
    f->sub(&t2, &t1, &vvv);
    f->small(&c2, 2);
    f->mul(&R2x, &c2, &R);
    f->sub(&A, &t2, &R2x);
    f->mul(&X3, &v, &A);
    f->sub(&RA, &R, &A);
    f->mul(&RA, &u, &RA);
    f->mul(&t1, &vvv, &Y1Z2);
    f->sub(&Y3, &RA, &t1);
    f->mul(&Z3, &vvv, &Z1Z2);
    memcpy(PX(r), &X3, f->sz);
    memcpy(PY(r, f), &Y3, f->sz);
    memcpy(PZ(r, f), &Z3, f->sz);
}

void or_g1_proj_add(OG1P* r, const OG1P* p1, const OG1P* p2) { proj_add(&F1, r, p1, p2); }
void or_g2_proj_add(OG2P* r, const OG2P* p1, const OG2P* p2) { proj_add(&F2, r, p1, p2); }
void or_g1_proj_double(OG1P* r, const OG1P* p) { proj_double(&F1, r, p); }
void or_g2_proj_double(OG2P* r, const OG2P* p) { proj_double(&F2, r, p); }

/* Curve equations actually satisfied by the reference generators:
 * G1: y^2 = x^3 + 1, G2: y^2 = x^3 + (1 + i). (The reference's curve_b_g1 = 4
 * and curve_b_g2 = 4i at src/curve.cu:280-312 do not hold for its generators
 * and are never read by the MSM formulas.) */
bool or_g1_is_on_curve(const OG1* p) {
    if (p->infinity) return true;
    OFp y2, x2, x3, one, rhs;
    or_fp_mont_mul(&y2, &p->y, &p->y);
    or_fp_mont_mul(&x2, &p->x, &p->x);
    or_fp_mont_mul(&x3, &x2, &p->x);
    or_fp_one_mont(&one);
    or_fp_add(&rhs, &x3, &one);
    return or_fp_cmp(&y2, &rhs) == 0;
}
bool or_g2_is_on_curve(const OG2* p) {
    if (p->infinity) return true;
    OFp2 y2, x2, x3, b, rhs;
    or_fp2_mont_mul(&y2, &p->y, &p->y);
    or_fp2_mont_mul(&x2, &p->x, &p->x);
    or_fp2_mont_mul(&x3, &x2, &p->x);
    or_fp_one_mont(&b.c0);
    b.c1 = b.c0;
    or_fp2_add(&rhs, &x3, &b);
    return or_fp_cmp(&y2.c0, &rhs.c0) == 0 && or_fp_cmp(&y2.c1, &rhs.c1) == 0;
}

/* Projective double-and-add, MSB first (the pattern of reference
 * src/curve.cu:231-268, in projective coordinates). */
static void scalar_mul_proj(const Field* f, void* out, const void* base_proj, const uint64_t* k, int limbs) {
    unsigned char acc[3 * sizeof(OFp2)];
    memset(acc, 0, sizeof acc);
    for (int i = limbs - 1; i >= 0; i--) {
        for (int b = 63; b >= 0; b--) {
            proj_double(f, acc, acc);
            if ((k[i] >> b) & 1) proj_add(f, acc, acc, base_proj);
        }
    }
    memcpy(out, acc, 3 * f->sz);
}

void or_g1_scalar_mul(OG1* out, const OG1* base, const uint64_t* k, int limbs) {
    OG1P bp, r;
    or_g1_affine_to_proj(&bp, base);
    scalar_mul_proj(&F1, &r, &bp, k, limbs);
    or_g1_proj_to_affine(out, &r);
}
void or_g2_scalar_mul(OG2* out, const OG2* base, const uint64_t* k, int limbs) {
    OG2P bp, r;
    or_g2_affine_to_proj(&bp, base);
    scalar_mul_proj(&F2, &r, &bp, k, limbs);
    or_g2_proj_to_affine(out, &r);
}

void or_g1_bases_from_k(OG1* out, const uint64_t* k, int n, int threads) {
    OG1 g;
    or_g1_generator(&g);
    or_g1_to_mont(&g, &g);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
    for (int i = 0; i < n; i++) or_g1_scalar_mul(&out[i], &g, &k[i], 1);
    (void)threads;
}
void or_g2_bases_from_k(OG2* out, const uint64_t* k, int n, int threads) {
    OG2 g;
    or_g2_generator(&g);
    or_g2_to_mont(&g, &g);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
#endif
    for (int i = 0; i < n; i++) or_g2_scalar_mul(&out[i], &g, &k[i], 1);
    (void)threads;
}

/* ------------------------------------------------------------------------ MSM */

/* Unsigned c-bit window w of a little-endian multi-limb scalar, including
 * windows that span two limbs (reference src/curve.cu:901-922). */
static uint32_t window_bits(const uint64_t* s, int limbs, int w, int c) {
    int off = w * c;
    int li = off / 64, bi = off % 64;
    if (li >= limbs) return 0;
    uint64_t v = s[li] >> bi;
    if (bi + c > 64 && li + 1 < limbs) v |= s[li + 1] << (64 - bi);
    return (uint32_t)(v & ((1ULL << c) - 1));
}

/* Bucket-method MSM (Pippenger): per window, bucket sums by projective adds,
 * window sum = sum_b b * B_b via a running sum, then Horner over windows from
 * the most significant down (the MSB->LSB combine of reference
 * src/curve.cu:2060-2179 / :2763-2820 with a general window width). */
static void msm_generic(const Field* f, void* out_proj, const void* bases_aff, size_t aff_stride,
                        void (*to_proj)(void*, const void*), const uint64_t* scalars, int limbs, int n, int c,
                        int threads) {
    size_t psz = 3 * f->sz;
    if (c <= 0) {
        c = 2;
        while (c < 16 && (1 << (c + 2)) < n) c++;
    }
    int nwin = (64 * limbs + c - 1) / c;
    int nb = 1 << c;
    unsigned char* wsum = (unsigned char*)calloc((size_t)nwin, psz);
    if (threads < 1) threads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int w = 0; w < nwin; w++) {
        unsigned char* buckets = (unsigned char*)calloc((size_t)nb, psz);
        unsigned char pt[3 * sizeof(OFp2)];
        for (int i = 0; i < n; i++) {
            uint32_t d = window_bits(scalars + (size_t)i * limbs, limbs, w, c);
            if (!d) continue;
            to_proj(pt, (const unsigned char*)bases_aff + (size_t)i * aff_stride);
            proj_add(f, buckets + d * psz, buckets + d * psz, pt);
        }
        unsigned char run[3 * sizeof(OFp2)], acc[3 * sizeof(OFp2)];
        memset(run, 0, sizeof run);
        memset(acc, 0, sizeof acc);
        for (int b = nb - 1; b >= 1; b--) {
            proj_add(f, run, run, buckets + b * psz);
            proj_add(f, acc, acc, run);
        }
        memcpy(wsum + (size_t)w * psz, acc, psz);
        free(buckets);
    }
    unsigned char res[3 * sizeof(OFp2)];
    memset(res, 0, sizeof res);
    for (int w = nwin - 1; w >= 0; w--) {
        for (int k = 0; k < c; k++) proj_double(f, res, res);
        proj_add(f, res, res, wsum + (size_t)w * psz);
    }
    memcpy(out_proj, res, psz);
    free(wsum);
}

static void g1_to_proj_v(void* o, const void* a) { or_g1_affine_to_proj((OG1P*)o, (const OG1*)a); }
static void g2_to_proj_v(void* o, const void* a) { or_g2_affine_to_proj((OG2P*)o, (const OG2*)a); }

void or_g1_msm(OG1* out, const OG1* bases, const uint64_t* scalars, int limbs, int n, int c, int threads) {
    OG1P r;
    msm_generic(&F1, &r, bases, sizeof(OG1), g1_to_proj_v, scalars, limbs, n, c, threads);
    or_g1_proj_to_affine(out, &r);
}
void or_g2_msm(OG2* out, const OG2* bases, const uint64_t* scalars, int limbs, int n, int c, int threads) {
    OG2P r;
    msm_generic(&F2, &r, bases, sizeof(OG2), g2_to_proj_v, scalars, limbs, n, c, threads);
    or_g2_proj_to_affine(out, &r);
}

/* ------------------------------------------------------------------------ RNG */

/* mt19937_64 (Matsumoto & Nishimura 2000; the C++ std::mt19937_64 engine). */
void or_mt64_seed(OMt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 312; i++)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->idx = 312;
}

uint64_t or_mt64_next(OMt64* s) {
    static const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL, A = 0xB5026F5AA96619E9ULL;
    if (s->idx >= 312) {
        for (int i = 0; i < 312; i++) {
            uint64_t x = (s->mt[i] & UM) | (s->mt[(i + 1) % 312] & LM);
            uint64_t xa = x >> 1;
            if (x & 1) xa ^= A;
            s->mt[i] = s->mt[(i + 156) % 312] ^ xa;
        }
        s->idx = 0;
    }
    uint64_t y = s->mt[s->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

void or_gen_u64(uint64_t seed, int n, uint64_t* out) {
    OMt64 s;
    or_mt64_seed(&s, seed);
    for (int i = 0; i < n; i++) out[i] = or_mt64_next(&s);
}

static int lt_r(const uint64_t* v) {
    for (int i = 4; i >= 0; i--) {
        if (v[i] < RORD[i]) return 1;
        if (v[i] > RORD[i]) return 0;
    }
    return 0;
}

void or_gen_scalars_modr(uint64_t seed, int n, uint64_t* out) {
    OMt64 s;
    or_mt64_seed(&s, seed);
    for (int i = 0; i < n; i++) {
        uint64_t v[5];
        do {
            for (int k = 0; k < 5; k++) v[k] = or_mt64_next(&s);
            v[4] &= (1ULL << 43) - 1;
        } while (!lt_r(v));
        memcpy(out + (size_t)i * 5, v, sizeof v);
    }
}

void or_scalar_order(uint64_t r_out[5]) { memcpy(r_out, RORD, sizeof RORD); }
