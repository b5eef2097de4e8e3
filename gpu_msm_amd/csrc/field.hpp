// Device field arithmetic for BLS12-446 on gfx950.
//
// Fp elements are 16 limbs of 28 bits held in 32-bit VGPRs, value in [0, 2p),
// Montgomery form with R = 2^448 (16 * 28 = 448, the same R as the reference's
// 7 x 64-bit representation, reference src/fp.cu:15-33, so the Montgomery
// images are identical and conversion is a pure bit repacking).
//
// Multiplication is product scanning with the Montgomery reduction folded into
// each column (FIPS). Every limb product is one v_mad_u64_u32 accumulating into
// a 64-bit column sum: 28-bit limbs keep a column of up to 32 products plus the
// incoming carry below 2^62, so no carry instruction is issued inside a
// column. On gfx950 v_mad_u64_u32 issues at full rate, so carry fix-ups would
// cost as much as the multiplies themselves.
//
// Lazy reduction: p < 2^446 so 4p < R; with inputs in [0, 2p) the Montgomery
// output stays in [0, 2p) without a final subtraction. add/sub return [0, 2p).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace gm {

constexpr int NL = 16;
constexpr uint32_t MASK28 = (1u << 28) - 1;

// p in 28-bit limbs (little endian)
constexpr uint32_t PL[NL] = {0xab0aaab, 0x1c0026a, 0x3b5cc31, 0x4528c57, 0xdee56ee, 0xc3e23ac,
                             0x2824e6d, 0xbac7160, 0xf75a64b, 0x02fe320, 0x5a4b78a, 0xd640009,
                             0x34965aa, 0x35200fc, 0xb28c5e5, 0x3cdee0f};
// 2p in 28-bit limbs
constexpr uint32_t P2L[NL] = {0x5615556, 0x38004d5, 0x76b9862, 0x8a518ae, 0xbdcaddc, 0x87c4759,
                              0x5049cdb, 0x758e2c0, 0xeeb4c97, 0x05fc641, 0xb496f14, 0xac80012,
                              0x692cb55, 0x6a401f8, 0x6518bca, 0x79bdc1f};
// -p^{-1} mod 2^28
constexpr uint32_t PINV28 = 0x035fffd;
// R mod p (Montgomery one)
constexpr uint32_t ONEL[NL] = {0x53d5554, 0x8fff655, 0x128cf3b, 0xeb5cea3, 0x846a446, 0xf07714c,
                               0x5f6c648, 0x14e3a7f, 0x22966d1, 0xf40737c, 0x96d21d7, 0xa6fffda,
                               0x2da6954, 0x2b7fc0f, 0x35ce86b, 0x0c847c1};
// R^2 mod p
constexpr uint32_t R2L[NL] = {0xc752b45, 0xff01ddd, 0xccf5b2a, 0x72a7421, 0xad692c7, 0x29c347d,
                              0xe140eef, 0x22ea595, 0x5a1400c, 0x145218f, 0x91c9fec, 0x3d4b99d,
                              0x37f9014, 0x4f3bb65, 0xc9be797, 0x3627854};

struct Fp {
    uint32_t v[NL];
};

struct Fp2 {
    Fp c0, c1;
};

__host__ __device__ __forceinline__ void fp_set_const(Fp& r, const uint32_t (&c)[NL]) {
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = c[i];
}
__host__ __device__ __forceinline__ void fp_set_zero(Fp& r) {
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = 0;
}
__host__ __device__ __forceinline__ void fp_set_one(Fp& r) { fp_set_const(r, ONEL); }

// r = a*b*R^{-1} mod p, inputs in [0, 2p), output in [0, 2p).
__device__ __forceinline__ void fp_mul(Fp& r, const Fp& a, const Fp& b) {
    uint32_t m[NL];
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) acc += (uint64_t)a.v[i] * b.v[k - i];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * PL[k - i];
        m[k] = ((uint32_t)acc * PINV28) & MASK28;
        acc += (uint64_t)m[k] * PL[0];
        acc >>= 28;
    }
#pragma unroll
    for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) acc += (uint64_t)a.v[i] * b.v[k - i];
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) acc += (uint64_t)m[i] * PL[k - i];
        r.v[k - NL] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    r.v[NL - 1] = (uint32_t)acc;
}

// r = a^2 * R^{-1}: off-diagonal products once against doubled limbs.
__device__ __forceinline__ void fp_sqr(Fp& r, const Fp& a) {
    uint32_t m[NL], d[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) d[i] = a.v[i] << 1;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
#pragma unroll
        for (int i = 0; i < k - i; ++i) acc += (uint64_t)d[i] * a.v[k - i];
        if ((k & 1) == 0) acc += (uint64_t)a.v[k >> 1] * a.v[k >> 1];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * PL[k - i];
        m[k] = ((uint32_t)acc * PINV28) & MASK28;
        acc += (uint64_t)m[k] * PL[0];
        acc >>= 28;
    }
#pragma unroll
    for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
        for (int i = k - NL + 1; i < k - i; ++i) acc += (uint64_t)d[i] * a.v[k - i];
        if ((k & 1) == 0) acc += (uint64_t)a.v[k >> 1] * a.v[k >> 1];
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) acc += (uint64_t)m[i] * PL[k - i];
        r.v[k - NL] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    r.v[NL - 1] = (uint32_t)acc;
}

// r = a + b mod p, [0, 2p) in and out.
__device__ __forceinline__ void fp_add(Fp& r, const Fp& a, const Fp& b) {
    uint32_t s[NL], d[NL];
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        uint32_t t = a.v[i] + b.v[i] + c;
        s[i] = t & MASK28;
        c = t >> 28;
    }
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        int32_t t = (int32_t)s[i] - (int32_t)P2L[i] + br;
        d[i] = (uint32_t)t & MASK28;
        br = t >> 28;
    }
    const bool keep = br < 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = keep ? s[i] : d[i];
}

// r = a - b mod p, [0, 2p) in and out.
__device__ __forceinline__ void fp_sub(Fp& r, const Fp& a, const Fp& b) {
    uint32_t d[NL];
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        int32_t t = (int32_t)a.v[i] - (int32_t)b.v[i] + br;
        d[i] = (uint32_t)t & MASK28;
        br = t >> 28;
    }
    const uint32_t msk = (uint32_t)br;  // all ones when a < b
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        uint32_t t = d[i] + (P2L[i] & msk) + c;
        r.v[i] = t & MASK28;
        c = t >> 28;
    }
}

__device__ __forceinline__ void fp_dbl(Fp& r, const Fp& a) { fp_add(r, a, a); }

// r = -a mod p for a in [0, 2p): 2p - a in (0, 2p], folded to [0, 2p).
__device__ __forceinline__ void fp_neg(Fp& r, const Fp& a) {
    Fp z;
    fp_set_zero(z);
    fp_sub(r, z, a);
}

// r = p - a for a canonical a in [0, p): in (0, p] (p, the other lazy form
// of 0, for a = 0), with normalised limbs, in one borrow chain.
__device__ __forceinline__ void fp_neg_canon(Fp& r, const Fp& a) {
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int32_t v = (int32_t)PL[i] - (int32_t)a.v[i] + br;
        r.v[i] = (uint32_t)v & MASK28;
        br = v >> 28;
    }
}

// Canonical representative in [0, p).
__device__ __forceinline__ void fp_canon(Fp& r, const Fp& a) {
    uint32_t d[NL];
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        int32_t t = (int32_t)a.v[i] - (int32_t)PL[i] + br;
        d[i] = (uint32_t)t & MASK28;
        br = t >> 28;
    }
    const bool keep = br < 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = keep ? a.v[i] : d[i];
}

// a == 0 mod p for a in [0, 2p): a is 0 or p.
__device__ __forceinline__ bool fp_is_zero(const Fp& a) {
    uint32_t z = 0, q = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        z |= a.v[i];
        q |= a.v[i] ^ PL[i];
    }
    return z == 0 || q == 0;
}

// Exact zero of the representation (used for the infinity marker ZZ = 0).
__device__ __forceinline__ bool fp_is_raw_zero(const Fp& a) {
    uint32_t z = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) z |= a.v[i];
    return z == 0;
}

__device__ __forceinline__ bool fp_eq(const Fp& a, const Fp& b) {
    Fp d;
    fp_sub(d, a, b);
    return fp_is_zero(d);
}

// acc = a^E in the Montgomery domain; E is nw 32-bit words, most significant
// first (left-to-right square and multiply).
__device__ inline void fp_pow_words(Fp& r, const Fp& a, const uint32_t* E, int nw) {
    Fp acc;
    fp_set_one(acc);
    for (int w = 0; w < nw; ++w) {
        const uint32_t e = E[w];
        for (int b = 31; b >= 0; --b) {
            fp_sqr(acc, acc);
            if ((e >> b) & 1) fp_mul(acc, acc, a);
        }
    }
    r = acc;
}

// p - 2 in 32-bit words, most significant first
constexpr uint32_t PM2_WORDS[14] = {0x3cdee0fb, 0x28c5e535, 0x200fc349, 0x65aad640, 0x0095a4b7,
                                    0x8a02fe32, 0x0f75a64b, 0xbac71602, 0x824e6dc3, 0xe23acdee,
                                    0x56ee4528, 0xc573b5cc, 0x311c0026, 0xaab0aaa9};

// (p + 1) / 4, (p - 1) / 2 and (p - 3) / 4 in 32-bit words, most significant
// first: square root (p = 3 mod 4), Euler's criterion and the Fp2 square
// root of Adj and Rodriguez-Henriquez (2012), Algorithm 9.
constexpr uint32_t SQRT_WORDS[14] = {0x0f37b83e, 0xca31794d, 0x4803f0d2, 0x596ab590, 0x0025692d,
                                     0xe280bf8c, 0x83dd6992, 0xeeb1c580, 0xa0939b70, 0xf88eb37b,
                                     0x95bb914a, 0x315ced73, 0x0c470009, 0xaaac2aab};
constexpr uint32_t EULER_WORDS[14] = {0x1e6f707d, 0x9462f29a, 0x9007e1a4, 0xb2d56b20, 0x004ad25b,
                                      0xc5017f19, 0x07bad325, 0xdd638b01, 0x412736e1, 0xf11d66f7,
                                      0x2b772294, 0x62b9dae6, 0x188e0013, 0x55585555};
constexpr uint32_t P34_WORDS[14] = {0x0f37b83e, 0xca31794d, 0x4803f0d2, 0x596ab590, 0x0025692d,
                                    0xe280bf8c, 0x83dd6992, 0xeeb1c580, 0xa0939b70, 0xf88eb37b,
                                    0x95bb914a, 0x315ced73, 0x0c470009, 0xaaac2aaa};
// (p - 1) / 2 in 28-bit limbs: canonical values above it are the "larger"
// of each pair {y, -y}.
constexpr uint32_t HALFL[NL] = {0x5585555, 0x8e00135, 0x9dae618, 0x229462b, 0x6f72b77, 0xe1f11d6,
                                0x1412736, 0xdd638b0, 0x7bad325, 0x017f190, 0xad25bc5, 0x6b20004,
                                0x1a4b2d5, 0x9a9007e, 0xd9462f2, 0x1e6f707};

// a^(p-2) (Fermat inverse); 0 maps to 0.
__device__ inline void fp_inv(Fp& r, const Fp& a) { fp_pow_words(r, a, PM2_WORDS, 14); }

// Column accumulation on the running column sum itself. In plain C++ the
// compiler sums each column's products on a chain of its own and adds it to
// the running value afterwards (one more 64-bit add per column); these keep
// the whole product on one chain. Both accumulation kernels take them (the
// G2 lane pairs through CHAIN = true, G1 through FpAccOps); the depth-bound
// kernels, whose few waves need the compiler's independent chains to hide
// the mad latency, keep plain C++.
#if defined(__HIP_DEVICE_COMPILE__)
// Up to eight mads on the running column sum in one asm statement: the
// compiler cannot see into an asm statement, so it puts a wait state
// (s_nop) between consecutive ones; inside one statement consecutive mads
// need none (a VGPR read after a VALU write interlocks in hardware, and the
// unused carry-outs only write vcc, which nothing reads). mad_ivN: signed
// x_j * y_j, both in VGPRs; mad_isN: y_j uniform constants in SGPRs (the
// reduction products m * p_j have non-negative operands below 2^31, so the
// signed mad equals the unsigned one). Arguments: x_0..x_{N-1}, y_0..y_{N-1}.
__device__ __forceinline__ void mad_iv1(int64_t& acc, int32_t x0, int32_t y0) {
    asm("v_mad_i64_i32 %0, vcc, %1, %2, %0" : "+v"(acc) : "v"(x0), "v"(y0) : "vcc");
}
__device__ __forceinline__ void mad_iv2(int64_t& acc, int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
    asm("v_mad_i64_i32 %0, vcc, %1, %3, %0\n\tv_mad_i64_i32 %0, vcc, %2, %4, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(y0), "v"(y1) : "vcc");
}
__device__ __forceinline__ void mad_iv3(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t y0, int32_t y1, int32_t y2) {
    asm("v_mad_i64_i32 %0, vcc, %1, %4, %0\n\tv_mad_i64_i32 %0, vcc, %2, %5, %0\n\tv_mad_i64_i32 %0, vcc, %3, %6, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(y0), "v"(y1), "v"(y2) : "vcc");
}
__device__ __forceinline__ void mad_iv4(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t y0, int32_t y1, int32_t y2, int32_t y3) {
    asm("v_mad_i64_i32 %0, vcc, %1, %5, %0\n\tv_mad_i64_i32 %0, vcc, %2, %6, %0\n\tv_mad_i64_i32 %0, vcc, %3, %7, %0\n\tv_mad_i64_i32 %0, vcc, %4, %8, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(y0), "v"(y1), "v"(y2), "v"(y3) : "vcc");
}
__device__ __forceinline__ void mad_iv5(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4) {
    asm("v_mad_i64_i32 %0, vcc, %1, %6, %0\n\tv_mad_i64_i32 %0, vcc, %2, %7, %0\n\tv_mad_i64_i32 %0, vcc, %3, %8, %0\n\tv_mad_i64_i32 %0, vcc, %4, %9, %0\n\tv_mad_i64_i32 %0, vcc, %5, %10, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(y0), "v"(y1), "v"(y2), "v"(y3), "v"(y4) : "vcc");
}
__device__ __forceinline__ void mad_iv6(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t x5, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4, int32_t y5) {
    asm("v_mad_i64_i32 %0, vcc, %1, %7, %0\n\tv_mad_i64_i32 %0, vcc, %2, %8, %0\n\tv_mad_i64_i32 %0, vcc, %3, %9, %0\n\tv_mad_i64_i32 %0, vcc, %4, %10, %0\n\tv_mad_i64_i32 %0, vcc, %5, %11, %0\n\tv_mad_i64_i32 %0, vcc, %6, %12, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(x5), "v"(y0), "v"(y1), "v"(y2), "v"(y3), "v"(y4), "v"(y5) : "vcc");
}
__device__ __forceinline__ void mad_iv7(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t x5, int32_t x6, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4, int32_t y5, int32_t y6) {
    asm("v_mad_i64_i32 %0, vcc, %1, %8, %0\n\tv_mad_i64_i32 %0, vcc, %2, %9, %0\n\tv_mad_i64_i32 %0, vcc, %3, %10, %0\n\tv_mad_i64_i32 %0, vcc, %4, %11, %0\n\tv_mad_i64_i32 %0, vcc, %5, %12, %0\n\tv_mad_i64_i32 %0, vcc, %6, %13, %0\n\tv_mad_i64_i32 %0, vcc, %7, %14, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(x5), "v"(x6), "v"(y0), "v"(y1), "v"(y2), "v"(y3), "v"(y4), "v"(y5), "v"(y6) : "vcc");
}
__device__ __forceinline__ void mad_iv8(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t x5, int32_t x6, int32_t x7, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4, int32_t y5, int32_t y6, int32_t y7) {
    asm("v_mad_i64_i32 %0, vcc, %1, %9, %0\n\tv_mad_i64_i32 %0, vcc, %2, %10, %0\n\tv_mad_i64_i32 %0, vcc, %3, %11, %0\n\tv_mad_i64_i32 %0, vcc, %4, %12, %0\n\tv_mad_i64_i32 %0, vcc, %5, %13, %0\n\tv_mad_i64_i32 %0, vcc, %6, %14, %0\n\tv_mad_i64_i32 %0, vcc, %7, %15, %0\n\tv_mad_i64_i32 %0, vcc, %8, %16, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(x5), "v"(x6), "v"(x7), "v"(y0), "v"(y1), "v"(y2), "v"(y3), "v"(y4), "v"(y5), "v"(y6), "v"(y7) : "vcc");
}
__device__ __forceinline__ void mad_is1(int64_t& acc, int32_t x0, int32_t y0) {
    asm("v_mad_i64_i32 %0, vcc, %1, %2, %0" : "+v"(acc) : "v"(x0), "s"(y0) : "vcc");
}
__device__ __forceinline__ void mad_is2(int64_t& acc, int32_t x0, int32_t x1, int32_t y0, int32_t y1) {
    asm("v_mad_i64_i32 %0, vcc, %1, %3, %0\n\tv_mad_i64_i32 %0, vcc, %2, %4, %0" : "+v"(acc) : "v"(x0), "v"(x1), "s"(y0), "s"(y1) : "vcc");
}
__device__ __forceinline__ void mad_is3(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t y0, int32_t y1, int32_t y2) {
    asm("v_mad_i64_i32 %0, vcc, %1, %4, %0\n\tv_mad_i64_i32 %0, vcc, %2, %5, %0\n\tv_mad_i64_i32 %0, vcc, %3, %6, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "s"(y0), "s"(y1), "s"(y2) : "vcc");
}
__device__ __forceinline__ void mad_is4(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t y0, int32_t y1, int32_t y2, int32_t y3) {
    asm("v_mad_i64_i32 %0, vcc, %1, %5, %0\n\tv_mad_i64_i32 %0, vcc, %2, %6, %0\n\tv_mad_i64_i32 %0, vcc, %3, %7, %0\n\tv_mad_i64_i32 %0, vcc, %4, %8, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "s"(y0), "s"(y1), "s"(y2), "s"(y3) : "vcc");
}
__device__ __forceinline__ void mad_is5(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4) {
    asm("v_mad_i64_i32 %0, vcc, %1, %6, %0\n\tv_mad_i64_i32 %0, vcc, %2, %7, %0\n\tv_mad_i64_i32 %0, vcc, %3, %8, %0\n\tv_mad_i64_i32 %0, vcc, %4, %9, %0\n\tv_mad_i64_i32 %0, vcc, %5, %10, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "s"(y0), "s"(y1), "s"(y2), "s"(y3), "s"(y4) : "vcc");
}
__device__ __forceinline__ void mad_is6(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t x5, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4, int32_t y5) {
    asm("v_mad_i64_i32 %0, vcc, %1, %7, %0\n\tv_mad_i64_i32 %0, vcc, %2, %8, %0\n\tv_mad_i64_i32 %0, vcc, %3, %9, %0\n\tv_mad_i64_i32 %0, vcc, %4, %10, %0\n\tv_mad_i64_i32 %0, vcc, %5, %11, %0\n\tv_mad_i64_i32 %0, vcc, %6, %12, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(x5), "s"(y0), "s"(y1), "s"(y2), "s"(y3), "s"(y4), "s"(y5) : "vcc");
}
__device__ __forceinline__ void mad_is7(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t x5, int32_t x6, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4, int32_t y5, int32_t y6) {
    asm("v_mad_i64_i32 %0, vcc, %1, %8, %0\n\tv_mad_i64_i32 %0, vcc, %2, %9, %0\n\tv_mad_i64_i32 %0, vcc, %3, %10, %0\n\tv_mad_i64_i32 %0, vcc, %4, %11, %0\n\tv_mad_i64_i32 %0, vcc, %5, %12, %0\n\tv_mad_i64_i32 %0, vcc, %6, %13, %0\n\tv_mad_i64_i32 %0, vcc, %7, %14, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(x5), "v"(x6), "s"(y0), "s"(y1), "s"(y2), "s"(y3), "s"(y4), "s"(y5), "s"(y6) : "vcc");
}
__device__ __forceinline__ void mad_is8(int64_t& acc, int32_t x0, int32_t x1, int32_t x2, int32_t x3, int32_t x4, int32_t x5, int32_t x6, int32_t x7, int32_t y0, int32_t y1, int32_t y2, int32_t y3, int32_t y4, int32_t y5, int32_t y6, int32_t y7) {
    asm("v_mad_i64_i32 %0, vcc, %1, %9, %0\n\tv_mad_i64_i32 %0, vcc, %2, %10, %0\n\tv_mad_i64_i32 %0, vcc, %3, %11, %0\n\tv_mad_i64_i32 %0, vcc, %4, %12, %0\n\tv_mad_i64_i32 %0, vcc, %5, %13, %0\n\tv_mad_i64_i32 %0, vcc, %6, %14, %0\n\tv_mad_i64_i32 %0, vcc, %7, %15, %0\n\tv_mad_i64_i32 %0, vcc, %8, %16, %0" : "+v"(acc) : "v"(x0), "v"(x1), "v"(x2), "v"(x3), "v"(x4), "v"(x5), "v"(x6), "v"(x7), "s"(y0), "s"(y1), "s"(y2), "s"(y3), "s"(y4), "s"(y5), "s"(y6), "s"(y7) : "vcc");
}

template <bool SGPR_Y, class... A>
__device__ __forceinline__ void mad_blk(int64_t& acc, A... a) {
    constexpr int N = (int)sizeof...(A) / 2;
    static_assert(N >= 1 && N <= 8, "one to eight mads per statement");
    if constexpr (SGPR_Y) {
        if constexpr (N == 1) mad_is1(acc, a...);
        else if constexpr (N == 2) mad_is2(acc, a...);
        else if constexpr (N == 3) mad_is3(acc, a...);
        else if constexpr (N == 4) mad_is4(acc, a...);
        else if constexpr (N == 5) mad_is5(acc, a...);
        else if constexpr (N == 6) mad_is6(acc, a...);
        else if constexpr (N == 7) mad_is7(acc, a...);
        else mad_is8(acc, a...);
    } else {
        if constexpr (N == 1) mad_iv1(acc, a...);
        else if constexpr (N == 2) mad_iv2(acc, a...);
        else if constexpr (N == 3) mad_iv3(acc, a...);
        else if constexpr (N == 4) mad_iv4(acc, a...);
        else if constexpr (N == 5) mad_iv5(acc, a...);
        else if constexpr (N == 6) mad_iv6(acc, a...);
        else if constexpr (N == 7) mad_iv7(acc, a...);
        else mad_iv8(acc, a...);
    }
}

// Column K of a product scan over T operand pairs (x_u, y_u): the limb
// products x_u[i] * y_u[K - i] for i from col_ilo(K) to col_ihi(K), numbered
// j = (i - col_ilo(K)) * T + u, issued in statements of up to eight.
constexpr int col_ilo(int K) { return K < NL ? 0 : K - NL + 1; }
constexpr int col_ihi(int K) { return K < NL ? K : NL - 1; }

template <int T, int K, int J0, size_t... Is>
__device__ __forceinline__ void col_prod_blk(int64_t& acc, const Fp* const (&x)[T],
                                             const int32_t (*const (&y)[T])[NL], std::index_sequence<Is...>) {
    mad_blk<false>(acc, (int32_t)x[(J0 + (int)Is) % T]->v[col_ilo(K) + (J0 + (int)Is) / T]...,
                   (*y[(J0 + (int)Is) % T])[K - col_ilo(K) - (J0 + (int)Is) / T]...);
}
template <int T, int K, int J0 = 0>
__device__ __forceinline__ void col_prods(int64_t& acc, const Fp* const (&x)[T], const int32_t (*const (&y)[T])[NL]) {
    constexpr int NT = (col_ihi(K) - col_ilo(K) + 1) * T;
    if constexpr (J0 < NT) {
        constexpr int N = NT - J0 < 8 ? NT - J0 : 8;
        col_prod_blk<T, K, J0>(acc, x, y, std::make_index_sequence<N>{});
        col_prods<T, K, J0 + N>(acc, x, y);
    }
}
// Reduction products m[i] * p[K - i] of column K for i in [I0, I1) (signed
// mads: an m[i] of -1 subtracts).
template <int K, int I0, int NM, size_t... Is>
__device__ __forceinline__ void col_red_blk(int64_t& acc, const uint32_t (&m)[NM], std::index_sequence<Is...>) {
    mad_blk<true>(acc, (int32_t)m[I0 + (int)Is]..., (int32_t)PL[K - I0 - (int)Is]...);
}
template <int K, int I0, int I1, int NM>
__device__ __forceinline__ void col_reds(int64_t& acc, const uint32_t (&m)[NM]) {
    if constexpr (I0 < I1) {
        constexpr int N = I1 - I0 < 8 ? I1 - I0 : 8;
        col_red_blk<K, I0>(acc, m, std::make_index_sequence<N>{});
        col_reds<K, I0 + N, I1>(acc, m);
    }
}
// The single-chain Montgomery product scan, one column per step. With OFFS,
// m[NL] (0 or -1, set by the caller) is one more row of the reduction at
// weight R = 2^448: the high columns take m[NL] * p[K - NL], which subtracts
// p R from the sum (the reduced value drops by p) where m[NL] = -1; its top
// limb p[NL - 1] is left to the caller.
template <int T, int K, bool OFFS = false>
__device__ __forceinline__ void redc_chain_cols(uint32_t (&t)[NL], int64_t& acc, uint32_t (&m)[NL + 1],
                                                const Fp* const (&x)[T], const int32_t (*const (&y)[T])[NL]) {
    if constexpr (K < 2 * NL - 1) {
        col_prods<T, K>(acc, x, y);
        if constexpr (K < NL) {
            col_reds<K, 0, K>(acc, m);
            m[K] = ((uint32_t)acc * PINV28) & MASK28;
            mad_blk<true>(acc, (int32_t)m[K], (int32_t)PL[0]);
        } else {
            col_reds<K, K - NL + 1, NL + (OFFS ? 1 : 0)>(acc, m);
            t[K - NL] = (uint32_t)acc & MASK28;
        }
        acc >>= 28;  // arithmetic: for K < NL the low 28 bits are zero
        redc_chain_cols<T, K + 1, OFFS>(t, acc, m, x, y);
    }
}

// Column K of a square of s (d = 2s): the off-diagonal products d[i] s[K - i]
// for i < K - i and, for even K, the diagonal s[K/2]^2, numbered j from 0
// (the diagonal last), issued in statements of up to eight.
template <int K>
struct SqCol {
    static constexpr int ilo = col_ilo(K);
    static constexpr int imax = (K + 1) / 2 - 1;  // largest i with i < K - i
    static constexpr int npair = imax >= ilo ? imax - ilo + 1 : 0;
    static constexpr int n = npair + (K % 2 == 0 ? 1 : 0);
};
template <int K, int J>
__device__ __forceinline__ int32_t sq_x(const int32_t (&d)[NL], const int32_t (&s)[NL]) {
    if constexpr (J < SqCol<K>::npair) return d[SqCol<K>::ilo + J];
    else return s[K / 2];
}
template <int K, int J>
__device__ __forceinline__ int32_t sq_y(const int32_t (&s)[NL]) {
    if constexpr (J < SqCol<K>::npair) return s[K - SqCol<K>::ilo - J];
    else return s[K / 2];
}
template <int K, int J0, size_t... Is>
__device__ __forceinline__ void sq_col_blk(int64_t& acc, const int32_t (&d)[NL], const int32_t (&s)[NL],
                                           std::index_sequence<Is...>) {
    mad_blk<false>(acc, sq_x<K, J0 + (int)Is>(d, s)..., sq_y<K, J0 + (int)Is>(s)...);
}
template <int K, int J0 = 0>
__device__ __forceinline__ void sq_col_prods(int64_t& acc, const int32_t (&d)[NL], const int32_t (&s)[NL]) {
    constexpr int NT = SqCol<K>::n;
    if constexpr (J0 < NT) {
        constexpr int N = NT - J0 < 8 ? NT - J0 : 8;
        sq_col_blk<K, J0>(acc, d, s, std::make_index_sequence<N>{});
        sq_col_prods<K, J0 + N>(acc, d, s);
    }
}
// The single-chain Montgomery square scan. With SUB, h[K - NL] joins column
// K of the high half (the fused X3 below); h[NL - 1] is left to the caller.
template <int K, bool SUB, int NM>
__device__ __forceinline__ void redc_sqr_chain_cols(uint32_t (&t)[NL], int64_t& acc, uint32_t (&m)[NM],
                                                    const int32_t (&d)[NL], const int32_t (&s)[NL],
                                                    const int32_t (&h)[NL]) {
    if constexpr (K < 2 * NL - 1) {
        sq_col_prods<K>(acc, d, s);
        if constexpr (K < NL) {
            col_reds<K, 0, K>(acc, m);
            m[K] = ((uint32_t)acc * PINV28) & MASK28;
            mad_blk<true>(acc, (int32_t)m[K], (int32_t)PL[0]);
        } else {
            col_reds<K, K - NL + 1, NL>(acc, m);
            if constexpr (SUB) acc += (int64_t)h[K - NL];
            t[K - NL] = (uint32_t)acc & MASK28;
        }
        acc >>= 28;
        redc_sqr_chain_cols<K + 1, SUB, NM>(t, acc, m, d, s, h);
    }
}
#endif

// Signed Montgomery reduction of a sum of limb products, the core of the
// lazily reduced products below: value = sum over the T terms of x_t * y_t,
// with y limbs signed (a negated operand makes a difference of products).
// Every product is one signed 32 x 32 -> 64 mad into the column sum; limbs
// are below 2^28 in magnitude, so T = 4 terms (64 products per column) plus
// 16 reduction products stay below 2^63. Returns the reduced value in
// 28-bit limbs with its top limb signed (in the last word, sign carried by
// the int64 returned as `top`). With OFFS, e * p R joins the sum (e is 0 or
// -1, per lane), so the reduced value drops by p where e = -1.
template <int T, bool CHAIN = false, bool OFFS = false>
__host__ __device__ __forceinline__ void redc_signed(uint32_t (&t)[NL], int64_t& top, const Fp* const (&x)[T],
                                            const int32_t (*const (&y)[T])[NL], int32_t e = 0) {
    uint32_t m[NL + 1];
    m[NL] = (uint32_t)e;
    int64_t acc = 0;
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (CHAIN) {
        // each column's limb products first, then its reduction products
        redc_chain_cols<T, 0, OFFS>(t, acc, m, x, y);
        top = acc;
        if constexpr (OFFS) top += (int64_t)e * (int32_t)PL[NL - 1];
        return;
    }
#endif
#pragma unroll
    for (int k = 0; k < NL; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i)
#pragma unroll
            for (int u = 0; u < T; ++u) acc += (int64_t)(int32_t)x[u]->v[i] * (*y[u])[k - i];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (int64_t)((uint64_t)m[i] * PL[k - i]);
        m[k] = ((uint32_t)acc * PINV28) & MASK28;
        acc += (int64_t)((uint64_t)m[k] * PL[0]);
        acc >>= 28;  // arithmetic: the low 28 bits are zero
    }
#pragma unroll
    for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) {
#pragma unroll
            for (int u = 0; u < T; ++u) acc += (int64_t)(int32_t)x[u]->v[i] * (*y[u])[k - i];
            acc += (int64_t)((uint64_t)m[i] * PL[k - i]);
        }
        if constexpr (OFFS) {
#if defined(__HIP_DEVICE_COMPILE__)
            // a mad into the column sum: the compiler would otherwise keep
            // every e * p[i] in registers ahead of its column
            mad_blk<true>(acc, e, (int32_t)PL[k - NL]);
#else
            acc += (int64_t)e * (int32_t)PL[k - NL];
#endif
        }
        t[k - NL] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    top = acc;
    if constexpr (OFFS) top += (int64_t)e * (int32_t)PL[NL - 1];
}

// Brings a reduced value in (-2p, 2p) (28-bit limbs 0..14 and a signed top
// limb) to [0, 2p): add 2p when negative. Values in (-2p, 3p) take
// fold_est<-1, 1> below.
__host__ __device__ __forceinline__ void redc_fold(Fp& r, uint32_t (&t)[NL], int64_t top) {
    const uint32_t msk = top < 0 ? 0xFFFFFFFFu : 0u;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        const int32_t v = (int32_t)t[i] + (int32_t)(P2L[i] & msk) + c;
        t[i] = (uint32_t)v & MASK28;
        c = v >> 28;
    }
    t[NL - 1] = (uint32_t)((int32_t)top + (int32_t)(P2L[NL - 1] & msk) + c);
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = t[i];
}

// r = (a b - c d) R^{-1} mod p with one reduction. Inputs in [0, 2p): the
// difference lies in (-4p^2, 4p^2), so the reduced value is in (-p, 2p) and
// one conditional addition of 2p lands in [0, 2p).
__host__ __device__ __forceinline__ void fp_mul_sub(Fp& r, const Fp& a, const Fp& b, const Fp& c, const Fp& d) {
    int32_t yb[NL], nd[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        yb[i] = (int32_t)b.v[i];
        nd[i] = -(int32_t)d.v[i];
    }
    const Fp* const x[2] = {&a, &c};
    const int32_t(*const y[2])[NL] = {&yb, &nd};
    uint32_t t[NL];
    int64_t top;
    redc_signed<2>(t, top, x, y);
    redc_fold(r, t, top);
}

// Raw differences. A value used only as a product operand can skip the
// carry normalisation and the 2p correction of fp_sub: r = a - b limb by limb
// (signed 32-bit limbs, |limb| < 2^28, value in (-2p, 2p) for a, b in
// [0, 2p)). The signed products below accept such operands: every limb
// product is a signed 32 x 32 -> 64 mad, and a product of two values below 2p
// in magnitude is below 4p^2, the same bound as for [0, 2p) operands.
__device__ __forceinline__ void fp_sub_raw(Fp& r, const Fp& a, const Fp& b) {
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = a.v[i] - b.v[i];
}

// r = a^2 R^{-1} for a raw a: the square is non-negative and below 4p^2, so
// the reduced value is in [0, 2p) with no correction (off-diagonal products
// once against doubled limbs, |2 a_i| < 2^29).
__device__ __forceinline__ void fp_sqr_raw(Fp& r, const Fp& a) {
    uint32_t m[NL];
    int32_t d[NL], s[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        s[i] = (int32_t)a.v[i];
        d[i] = s[i] * 2;
    }
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
#pragma unroll
        for (int i = 0; i < k - i; ++i) acc += (int64_t)d[i] * s[k - i];
        if ((k & 1) == 0) acc += (int64_t)s[k >> 1] * s[k >> 1];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (int64_t)((uint64_t)m[i] * PL[k - i]);
        m[k] = ((uint32_t)acc * PINV28) & MASK28;
        acc += (int64_t)((uint64_t)m[k] * PL[0]);
        acc >>= 28;  // arithmetic: the low 28 bits are zero
    }
#pragma unroll
    for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
        for (int i = k - NL + 1; i < k - i; ++i) acc += (int64_t)d[i] * s[k - i];
        if ((k & 1) == 0) acc += (int64_t)s[k >> 1] * s[k >> 1];
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) acc += (int64_t)((uint64_t)m[i] * PL[k - i]);
        r.v[k - NL] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    r.v[NL - 1] = (uint32_t)acc;  // the value is in [0, 2p): the top limb is non-negative
}

// Subtracts m (4p or 2p limbs) from the normalised value t when that does not
// borrow.
__host__ __device__ __forceinline__ void fp_cond_sub(uint32_t (&t)[NL], const uint32_t (&m)[NL]) {
    uint32_t d[NL];
    int32_t br = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int32_t v = (int32_t)t[i] - (int32_t)m[i] + br;
        d[i] = (uint32_t)v & MASK28;
        br = v >> 28;
    }
    const bool keep = br < 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) t[i] = keep ? t[i] : d[i];
}

// floor(j 2p / 2^420) for j >= 0: limb 15 of j 2p with the carries of its
// lower limbs.
constexpr int32_t mul2p_top(int j) {
    int64_t c = 0;
    for (int i = 0; i < NL - 1; ++i) c = ((int64_t)j * P2L[i] + c) >> 28;
    return (int32_t)((int64_t)j * P2L[NL - 1] + c);
}
// floor(j 2p / 2^420) for any small j (for j != 0 the low 420 bits of j 2p
// are not zero: the lowest limb of 2p is not).
constexpr int32_t mul2p_floor_top(int j) { return j >= 0 ? mul2p_top(j) : -mul2p_top(-j) - 1; }

// Brings a value v in [KMIN 2p, (KMAX + 1) 2p) to [0, 2p) in one pass: v has
// normalised 28-bit limbs 0..14 and a signed limb 15 `top` (bits 420 and up).
// The quotient k = floor(v / 2p) is the number of multiples j 2p
// (KMIN < j <= KMAX) at or below v, less -KMIN, and each comparison needs only
// top: v >= j 2p when top > floor(j 2p / 2^420), v < j 2p when top is smaller
// (for j = 0 the sign of top decides). Equal top limbs leave the comparison
// open; counting it as below makes k at most one short. v - k 2p (limbs
// t_i - k 2p_i with carries) then lies in [0, 2p), or in [0, 4p) after a tie,
// for which one conditional subtraction of 2p follows. A tie needs the top
// 26 bits of v to equal those of a multiple of 2p: about once in 10^8 values,
// so the subtraction is a branch the whole wave skips almost always.
template <int KMIN, int KMAX>
__host__ __device__ __forceinline__ void fold_est(Fp& r, uint32_t (&t)[NL], int32_t top) {
    int32_t k = KMIN;
    bool tie = false;
#pragma unroll
    for (int j = KMIN + 1; j <= KMAX; ++j) {
        if (j == 0) {
            k += top >= 0 ? 1 : 0;
        } else {
            const int32_t tj = mul2p_floor_top(j);
            k += top > tj ? 1 : 0;
            tie = tie || top == tj;
        }
    }
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        const int32_t v = (int32_t)t[i] - k * (int32_t)P2L[i] + c;
        t[i] = (uint32_t)v & MASK28;
        c = v >> 28;
    }
    t[NL - 1] = (uint32_t)(top - k * (int32_t)P2L[NL - 1] + c);
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_amdgcn_ballot_w64(tie) != 0) fp_cond_sub(t, P2L);
#else
    if (tie) fp_cond_sub(t, P2L);
#endif
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = t[i];
}

// r = a - b - 2c for a, b, c in [0, 2p), in [0, 2p): a - b - 2c + 6p lies in
// (0, 8p), summed limb by limb with the carries (limb i of 6p - b - 2c is in
// (-3 * 2^28, 6 * 2^28)), then one fold pass. The X3 = RR - PPP - 2Q and
// X3 = M^2 - 2S of the split point formulas, in two passes instead of the
// four or six of separate subtractions.
__device__ __forceinline__ void fp_sub_b2c(Fp& r, const Fp& a, const Fp& b, const Fp& c) {
    uint32_t t[NL];
    int32_t acc = 0;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        acc += (int32_t)a.v[i] - (int32_t)b.v[i] - 2 * (int32_t)c.v[i] + (int32_t)(6u * PL[i]);
        t[i] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    acc += (int32_t)a.v[NL - 1] - (int32_t)b.v[NL - 1] - 2 * (int32_t)c.v[NL - 1] + (int32_t)(6u * PL[NL - 1]);
    fold_est<0, 3>(r, t, acc);
}

// r = 3a for a in [0, 2p), in [0, 2p): 3a lies in [0, 6p), tripled limb by
// limb with the carries, then one fold pass.
__device__ __forceinline__ void fp_triple(Fp& r, const Fp& a) {
    uint32_t t[NL];
    int32_t acc = 0;
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) {
        acc += 3 * (int32_t)a.v[i];
        t[i] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    acc += 3 * (int32_t)a.v[NL - 1];
    fold_est<0, 2>(r, t, acc);
}

// Brings a reduced value in (-4p, 4p) (28-bit limbs 0..14 and a signed top
// limb) to [0, 2p), in one pass (fold_est).
__device__ __forceinline__ void redc_fold_wide(Fp& r, uint32_t (&t)[NL], int64_t top) {
    fold_est<-2, 1>(r, t, (int32_t)top);
}

// r = a^2 R^{-1} - b - 2c for a raw a and b, c in [0, 2p), in [0, 2p) (the
// X3 of every XYZZ formula). b + 2c - 6p is subtracted from the high half of
// the square's column sums, i.e. (b + 2c - 6p) R before the reduction: the
// square reduces to [0, 1.95p) and b + 2c < 6p, so the result lies in
// (0, 8p) with normalised limbs, and one quotient-estimate pass
// (fold_est<0, 3>) brings it to [0, 2p). Limb i of 6p - b - 2c is
// 6 p_i - b_i - 2 c_i, in (-3 * 2^28, 2^31).
__device__ __forceinline__ void fp_sqr_raw_sub2(Fp& r, const Fp& a, const Fp& b, const Fp& c) {
    uint32_t m[NL], t[NL];
    int32_t d[NL], s[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        s[i] = (int32_t)a.v[i];
        d[i] = s[i] * 2;
    }
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
#pragma unroll
        for (int i = 0; i < k - i; ++i) acc += (int64_t)d[i] * s[k - i];
        if ((k & 1) == 0) acc += (int64_t)s[k >> 1] * s[k >> 1];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (int64_t)((uint64_t)m[i] * PL[k - i]);
        m[k] = ((uint32_t)acc * PINV28) & MASK28;
        acc += (int64_t)((uint64_t)m[k] * PL[0]);
        acc >>= 28;
    }
#pragma unroll
    for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
        for (int i = k - NL + 1; i < k - i; ++i) acc += (int64_t)d[i] * s[k - i];
        if ((k & 1) == 0) acc += (int64_t)s[k >> 1] * s[k >> 1];
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) acc += (int64_t)((uint64_t)m[i] * PL[k - i]);
        acc += (int64_t)((int32_t)(6u * PL[k - NL]) - (int32_t)b.v[k - NL] - 2 * (int32_t)c.v[k - NL]);
        t[k - NL] = (uint32_t)acc & MASK28;
        acc >>= 28;
    }
    acc += (int64_t)((int32_t)(6u * PL[NL - 1]) - (int32_t)b.v[NL - 1] - 2 * (int32_t)c.v[NL - 1]);
    fold_est<0, 3>(r, t, (int32_t)acc);  // in (0, 8p): non-negative, below 2^29
}

// r = a b R^{-1} for a in [0, 2p) and a raw b: the product lies in
// (-4p^2, 4p^2), the reduced value in (-p, 2p), and one conditional addition
// of 2p lands in [0, 2p).
__device__ __forceinline__ void fp_mul_raw(Fp& r, const Fp& a, const Fp& b) {
    const Fp* const x[1] = {&a};
    const int32_t(*const y[1])[NL] = {reinterpret_cast<const int32_t(*)[NL]>(&b.v)};
    uint32_t t[NL];
    int64_t top;
    redc_signed<1>(t, top, x, y);
    redc_fold(r, t, top);
}

// r = (a b - c d) R^{-1} for raw a, b and c, d in [0, 2p): a b lies in
// (-4p^2, 4p^2) and c d in [0, 4p^2), so the difference lies in (-8p^2, 4p^2)
// and the reduced value in (-1.9p, 1.95p) (8p^2 / R < 1.9p as p < 0.95 R / 4);
// one conditional addition of 2p lands in [0, 2p).
__device__ __forceinline__ void fp_mul_sub_raw(Fp& r, const Fp& a, const Fp& b, const Fp& c, const Fp& d) {
    int32_t nd[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) nd[i] = -(int32_t)d.v[i];
    const Fp* const x[2] = {&a, &c};
    const int32_t(*const y[2])[NL] = {reinterpret_cast<const int32_t(*)[NL]>(&b.v), &nd};
    uint32_t t[NL];
    int64_t top;
    redc_signed<2>(t, top, x, y);
    redc_fold(r, t, top);
}

// ------------------------------------------------------------------ Fp2
// Fp2 = Fp[i]/(i^2 + 1) (non-residue -1 as in reference include/fp2.h:7).

__device__ __forceinline__ void fp2_add(Fp2& r, const Fp2& a, const Fp2& b) {
    fp_add(r.c0, a.c0, b.c0);
    fp_add(r.c1, a.c1, b.c1);
}
__device__ __forceinline__ void fp2_sub(Fp2& r, const Fp2& a, const Fp2& b) {
    fp_sub(r.c0, a.c0, b.c0);
    fp_sub(r.c1, a.c1, b.c1);
}
__device__ __forceinline__ void fp2_dbl(Fp2& r, const Fp2& a) { fp2_add(r, a, a); }
__device__ __forceinline__ void fp2_neg(Fp2& r, const Fp2& a) {
    fp_neg(r.c0, a.c0);
    fp_neg(r.c1, a.c1);
}
// (a0 + a1 i)(b0 + b1 i) = (a0 b0 - a1 b1) + (a0 b1 + a1 b0) i with both
// components reduced lazily: the two double-width products of each component
// are accumulated into one column sum before a single Montgomery reduction,
// so the product costs two reductions instead of three (Karatsuba over three
// fully reduced Fp products) and no intermediate Fp add/sub.
//   c0 columns are signed: a1 b1 enters through signed mads with -b1. With
//   inputs in [0, 2p), |a0 b0 - a1 b1| < 4p^2, so the reduced value lies in
//   (-4p^2/R, 4p^2/R + p) within (-p, 2p); a negative result gets 2p added.
//   c1 = a0 b1 + a1 b0 < 8p^2 reduces to below 8p^2/R + p < 3p; one
//   conditional subtraction of 2p returns it to [0, 2p).
// Column bounds: limbs < 2^28, so each product < 2^56; 32 products plus 16
// reduction terms stay below 2^62 in magnitude.
__host__ __device__ __forceinline__ void fp2_mul(Fp2& r, const Fp2& a, const Fp2& b) {
    int32_t nb1[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) nb1[i] = -(int32_t)b.c1.v[i];
    uint32_t m0[NL], m1[NL];
    int64_t acc0 = 0;
    uint64_t acc1 = 0;
#pragma unroll
    for (int k = 0; k < NL; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) {
            acc0 += (int64_t)(int32_t)a.c0.v[i] * (int32_t)b.c0.v[k - i];
            acc0 += (int64_t)(int32_t)a.c1.v[i] * nb1[k - i];
            acc1 += (uint64_t)a.c0.v[i] * b.c1.v[k - i];
            acc1 += (uint64_t)a.c1.v[i] * b.c0.v[k - i];
        }
#pragma unroll
        for (int i = 0; i < k; ++i) {
            acc0 += (int64_t)((uint64_t)m0[i] * PL[k - i]);
            acc1 += (uint64_t)m1[i] * PL[k - i];
        }
        m0[k] = ((uint32_t)acc0 * PINV28) & MASK28;
        m1[k] = ((uint32_t)acc1 * PINV28) & MASK28;
        acc0 += (int64_t)((uint64_t)m0[k] * PL[0]);
        acc1 += (uint64_t)m1[k] * PL[0];
        acc0 >>= 28;  // arithmetic: the low 28 bits are zero
        acc1 >>= 28;
    }
    uint32_t r0[NL], r1[NL];
#pragma unroll
    for (int k = NL; k < 2 * NL - 1; ++k) {
#pragma unroll
        for (int i = k - NL + 1; i < NL; ++i) {
            acc0 += (int64_t)(int32_t)a.c0.v[i] * (int32_t)b.c0.v[k - i];
            acc0 += (int64_t)(int32_t)a.c1.v[i] * nb1[k - i];
            acc1 += (uint64_t)a.c0.v[i] * b.c1.v[k - i];
            acc1 += (uint64_t)a.c1.v[i] * b.c0.v[k - i];
            acc0 += (int64_t)((uint64_t)m0[i] * PL[k - i]);
            acc1 += (uint64_t)m1[i] * PL[k - i];
        }
        r0[k - NL] = (uint32_t)acc0 & MASK28;
        r1[k - NL] = (uint32_t)acc1 & MASK28;
        acc0 >>= 28;
        acc1 >>= 28;
    }
    // c0: top limb carries the sign; add 2p when negative.
    {
        const uint32_t msk = acc0 < 0 ? 0xFFFFFFFFu : 0u;
        int32_t c = 0;
#pragma unroll
        for (int i = 0; i < NL - 1; ++i) {
            const int32_t t = (int32_t)r0[i] + (int32_t)(P2L[i] & msk) + c;
            r.c0.v[i] = (uint32_t)t & MASK28;
            c = t >> 28;
        }
        r.c0.v[NL - 1] = (uint32_t)((int32_t)acc0 + (int32_t)(P2L[NL - 1] & msk) + c);
    }
    // c1: subtract 2p unless that borrows.
    {
        r1[NL - 1] = (uint32_t)acc1;
        uint32_t d[NL];
        int32_t br = 0;
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int32_t t = (int32_t)r1[i] - (int32_t)P2L[i] + br;
            d[i] = (uint32_t)t & MASK28;
            br = t >> 28;
        }
        const bool keep = br < 0;
#pragma unroll
        for (int i = 0; i < NL; ++i) r.c1.v[i] = keep ? r1[i] : d[i];
    }
}
// (a0 + a1 i)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 i
__device__ __forceinline__ void fp2_sqr(Fp2& r, const Fp2& a) {
    Fp s, d, t;
    fp_add(s, a.c0, a.c1);
    fp_sub(d, a.c0, a.c1);
    fp_mul(t, a.c0, a.c1);
    fp_mul(r.c0, s, d);
    fp_dbl(r.c1, t);
}
__device__ __forceinline__ bool fp2_is_zero(const Fp2& a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
__device__ __forceinline__ bool fp2_is_raw_zero(const Fp2& a) {
    return fp_is_raw_zero(a.c0) && fp_is_raw_zero(a.c1);
}
__device__ __forceinline__ void fp2_canon(Fp2& r, const Fp2& a) {
    fp_canon(r.c0, a.c0);
    fp_canon(r.c1, a.c1);
}
__device__ __forceinline__ void fp2_inv(Fp2& r, const Fp2& a) {
    Fp t0, t1, n;
    fp_sqr(t0, a.c0);
    fp_sqr(t1, a.c1);
    fp_add(n, t0, t1);
    fp_inv(n, n);
    fp_mul(r.c0, a.c0, n);
    fp_neg(t0, a.c1);
    fp_mul(r.c1, t0, n);
}

// Field policies so the curve code is written once for G1 (Fp) and G2 (Fp2).
struct FpOps {
    using F = Fp;
    static constexpr int WORDS = NL;
    static constexpr bool RAW = true;  // has sqr_sub2 / sqr_sub2z
    __device__ static __forceinline__ void add(F& r, const F& a, const F& b) { fp_add(r, a, b); }
    __device__ static __forceinline__ void sub(F& r, const F& a, const F& b) { fp_sub(r, a, b); }
    __device__ static __forceinline__ void dbl(F& r, const F& a) { fp_dbl(r, a); }
    __device__ static __forceinline__ void neg(F& r, const F& a) { fp_neg(r, a); }
    // -a for a canonical a (the resident bases' coordinates)
    __device__ static __forceinline__ void neg_canon(F& r, const F& a) { fp_neg_canon(r, a); }
    __device__ static __forceinline__ void mul(F& r, const F& a, const F& b) { fp_mul(r, a, b); }
    __device__ static __forceinline__ void sqr(F& r, const F& a) { fp_sqr(r, a); }
    // r = a b - c d
    __device__ static __forceinline__ void mul_sub(F& r, const F& a, const F& b, const F& c, const F& d) {
        fp_mul_sub(r, a, b, c, d);
    }
    // Raw differences (operands of products only) and the products taking
    // them: sqr_raw(raw), mul_raw(normalised, raw), mul_sub_raw(raw, raw,
    // normalised, normalised); all results are normalised to [0, 2p).
    __device__ static __forceinline__ void sub_raw(F& r, const F& a, const F& b) { fp_sub_raw(r, a, b); }
    __device__ static __forceinline__ void sqr_raw(F& r, const F& a) { fp_sqr_raw(r, a); }
    __device__ static __forceinline__ void mul_raw(F& r, const F& a, const F& b) { fp_mul_raw(r, a, b); }
    __device__ static __forceinline__ void mul_sub_raw(F& r, const F& a, const F& b, const F& c, const F& d) {
        fp_mul_sub_raw(r, a, b, c, d);
    }
    // r = a^2 - b - 2c and r = a^2 - 2c (a raw)
    __device__ static __forceinline__ void sqr_sub2(F& r, const F& a, const F& b, const F& c) {
        fp_sqr_raw_sub2(r, a, b, c);
    }
    __device__ static __forceinline__ void sqr_sub2z(F& r, const F& a, const F& c) {
        Fp z;
        fp_set_zero(z);
        fp_sqr_raw_sub2(r, a, z, c);
    }
    __device__ static __forceinline__ bool is_zero(const F& a) { return fp_is_zero(a); }
    __device__ static __forceinline__ bool is_raw_zero(const F& a) { return fp_is_raw_zero(a); }
    __device__ static __forceinline__ void canon(F& r, const F& a) { fp_canon(r, a); }
    __device__ static __forceinline__ void inv(F& r, const F& a) { fp_inv(r, a); }
    __device__ static __forceinline__ void zero(F& r) { fp_set_zero(r); }
    __device__ static __forceinline__ void one(F& r) { fp_set_one(r); }
};

// Single-chain forms of the G1 products of the mixed addition: the same
// values and bounds as fp_mul, fp_mul_raw, fp_mul_sub_raw, fp_sqr_raw and
// fp_sqr_raw_sub2, every column summed on the running column sum itself.
// The host pass (which never runs them) compiles the plain forms.
__device__ __forceinline__ void fp_mul_chain(Fp& r, const Fp& a, const Fp& b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const Fp* const x[1] = {&a};
    const int32_t(*const y[1])[NL] = {reinterpret_cast<const int32_t(*)[NL]>(&b.v)};
    uint32_t t[NL];
    int64_t top;
    redc_signed<1, true>(t, top, x, y);
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) r.v[i] = t[i];
    r.v[NL - 1] = (uint32_t)top;  // a b < 4p^2 reduces to [0, 2p)
#else
    fp_mul(r, a, b);
#endif
}
__device__ __forceinline__ void fp_mul_raw_chain(Fp& r, const Fp& a, const Fp& b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const Fp* const x[1] = {&a};
    const int32_t(*const y[1])[NL] = {reinterpret_cast<const int32_t(*)[NL]>(&b.v)};
    uint32_t t[NL];
    int64_t top;
    redc_signed<1, true>(t, top, x, y);
    redc_fold(r, t, top);
#else
    fp_mul_raw(r, a, b);
#endif
}
__device__ __forceinline__ void fp_mul_sub_raw_chain(Fp& r, const Fp& a, const Fp& b, const Fp& c, const Fp& d) {
#if defined(__HIP_DEVICE_COMPILE__)
    int32_t nd[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) nd[i] = -(int32_t)d.v[i];
    const Fp* const x[2] = {&a, &c};
    const int32_t(*const y[2])[NL] = {reinterpret_cast<const int32_t(*)[NL]>(&b.v), &nd};
    uint32_t t[NL];
    int64_t top;
    redc_signed<2, true>(t, top, x, y);
    redc_fold(r, t, top);
#else
    fp_mul_sub_raw(r, a, b, c, d);
#endif
}
__device__ __forceinline__ void fp_sqr_raw_chain(Fp& r, const Fp& a) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t m[NL], t[NL];
    int32_t d[NL], s[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        s[i] = (int32_t)a.v[i];
        d[i] = s[i] * 2;
    }
    int64_t acc = 0;
    redc_sqr_chain_cols<0, false>(t, acc, m, d, s, s);
#pragma unroll
    for (int i = 0; i < NL - 1; ++i) r.v[i] = t[i];
    r.v[NL - 1] = (uint32_t)acc;  // the square is in [0, 2p): the top limb is non-negative
#else
    fp_sqr_raw(r, a);
#endif
}
// a^2 - b - 2c for a raw a (fp_sqr_raw_sub2: (0, 8p) before the fold).
__device__ __forceinline__ void fp_sqr_raw_sub2_chain(Fp& r, const Fp& a, const Fp& b, const Fp& c) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t m[NL], t[NL];
    int32_t d[NL], s[NL], h[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        s[i] = (int32_t)a.v[i];
        d[i] = s[i] * 2;
        h[i] = (int32_t)(6u * PL[i]) - (int32_t)b.v[i] - 2 * (int32_t)c.v[i];
    }
    int64_t acc = 0;
    redc_sqr_chain_cols<0, true>(t, acc, m, d, s, h);
    acc += (int64_t)h[NL - 1];
    fold_est<0, 3>(r, t, (int32_t)acc);  // in (0, 8p): non-negative, below 2^29
#else
    fp_sqr_raw_sub2(r, a, b, c);
#endif
}

// G1 policy of the accumulation kernel: FpOps with the single-chain products
// (the kernel runs at two waves per SIMD, whose second wave hides the longer
// dependency chain).
struct FpAccOps : FpOps {
    __device__ static __forceinline__ void mul(F& r, const F& a, const F& b) { fp_mul_chain(r, a, b); }
    __device__ static __forceinline__ void sqr_raw(F& r, const F& a) { fp_sqr_raw_chain(r, a); }
    __device__ static __forceinline__ void mul_raw(F& r, const F& a, const F& b) { fp_mul_raw_chain(r, a, b); }
    __device__ static __forceinline__ void mul_sub_raw(F& r, const F& a, const F& b, const F& c, const F& d) {
        fp_mul_sub_raw_chain(r, a, b, c, d);
    }
    __device__ static __forceinline__ void sqr_sub2(F& r, const F& a, const F& b, const F& c) {
        fp_sqr_raw_sub2_chain(r, a, b, c);
    }
    __device__ static __forceinline__ void sqr_sub2z(F& r, const F& a, const F& c) {
        Fp z;
        fp_set_zero(z);
        fp_sqr_raw_sub2_chain(r, a, z, c);
    }
};

struct Fp2Ops {
    using F = Fp2;
    static constexpr int WORDS = 2 * NL;
    static constexpr bool RAW = false;
    __device__ static __forceinline__ void add(F& r, const F& a, const F& b) { fp2_add(r, a, b); }
    __device__ static __forceinline__ void sub(F& r, const F& a, const F& b) { fp2_sub(r, a, b); }
    __device__ static __forceinline__ void dbl(F& r, const F& a) { fp2_dbl(r, a); }
    __device__ static __forceinline__ void neg(F& r, const F& a) { fp2_neg(r, a); }
    __device__ static __forceinline__ void mul(F& r, const F& a, const F& b) { fp2_mul(r, a, b); }
    __device__ static __forceinline__ void sqr(F& r, const F& a) { fp2_sqr(r, a); }
    __device__ static __forceinline__ void mul_sub(F& r, const F& a, const F& b, const F& c, const F& d) {
        F t;
        fp2_mul(t, c, d);
        fp2_mul(r, a, b);
        fp2_sub(r, r, t);
    }
    // no raw form: the "raw" operations are the normalised ones
    __device__ static __forceinline__ void sub_raw(F& r, const F& a, const F& b) { sub(r, a, b); }
    __device__ static __forceinline__ void sqr_raw(F& r, const F& a) { sqr(r, a); }
    __device__ static __forceinline__ void mul_raw(F& r, const F& a, const F& b) { mul(r, a, b); }
    __device__ static __forceinline__ void mul_sub_raw(F& r, const F& a, const F& b, const F& c, const F& d) {
        mul_sub(r, a, b, c, d);
    }
    __device__ static __forceinline__ bool is_zero(const F& a) { return fp2_is_zero(a); }
    __device__ static __forceinline__ bool is_raw_zero(const F& a) { return fp2_is_raw_zero(a); }
    __device__ static __forceinline__ void canon(F& r, const F& a) { fp2_canon(r, a); }
    __device__ static __forceinline__ void inv(F& r, const F& a) { fp2_inv(r, a); }
    __device__ static __forceinline__ void zero(F& r) {
        fp_set_zero(r.c0);
        fp_set_zero(r.c1);
    }
    __device__ static __forceinline__ void one(F& r) {
        fp_set_one(r.c0);
        fp_set_zero(r.c1);
    }
};

// ------------------------------------------------- Fp2 over a lane pair
// G2 alternative to Fp2Ops for the bucket kernels: lanes 2k and 2k+1 of a
// wave carry components c0 and c1 of the same Fp2 values, so each lane holds
// half the state of a point (a G2 XYZZ point is 64 registers per lane instead
// of 128), which keeps the mixed addition below 256 registers and two waves
// per SIMD in flight. A product reads the partner's component of one operand
// and both components of the other through DPP quad permutations, and each
// lane reduces its own component:
//   part 0: a0 b0 - a1 b1,   part 1: a1 b0 + a0 b1,
// both as x y + u w with signed w (w = -b1 for part 0, +b1 for part 1), one
// fused column sum and one Montgomery reduction per lane; the result is
// brought into [0, 2p) by a conditional addition of 2p (fp2_lane_mul) or,
// for raw operands and differences of products, by corrections in both
// directions. Every branch that depends on a value
// combines both lanes first, so control flow is uniform over each pair and
// the DPP partner is always active.

__device__ __forceinline__ uint32_t lane_part() { return threadIdx.x & 1u; }

// Value of v in the partner lane (quad_perm [1, 0, 3, 2]).
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ void fp_pair_swap(Fp& r, const Fp& a) {
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = pair_swap(a.v[i]);
}

// Value of v in the pair's part-0 lane (quad_perm [0, 0, 2, 2]) and in its
// part-1 lane ([1, 1, 3, 3]), given to both lanes of the pair. pair_part1
// is written as an update with 0 as the old value, so the compiler folds
// the read into a following xor or add (v_xor_b32_dpp, v_add_u32_dpp).
__device__ __forceinline__ uint32_t pair_part0(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xA0, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t pair_part1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xF5, 0xF, 0xF, true);
}

// Signed limbs of the operands a lane multiplies for its own component of
// (a0 + a1 i)(b0 + b1 i) = (a0 b0 - a1 b1) + (a1 b0 + a0 b1) i, as
// a y + ao w with ao the partner's part of a: y = b0 in both lanes, w = -b1
// in part 0 and +b1 in part 1, each read from its lane of the pair. The sign
// is (v ^ s) - s with s = -1 where w is negated. `neg` negates both (for a
// subtracted product): y = -b0, w = +b1 in part 0 and -b1 in part 1.
__device__ __forceinline__ void lane_operands(int32_t (&y)[NL], int32_t (&w)[NL], const Fp& b, bool neg) {
    int32_t s = ((lane_part() != 0) == neg) ? -1 : 0;
    asm("" : "+v"(s));  // keeps the xor and subtract, where a select would not take the DPP read
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int32_t b0 = (int32_t)pair_part0(b.v[i]);
        y[i] = neg ? -b0 : b0;
        w[i] = ((int32_t)pair_part1(b.v[i]) ^ s) - s;
    }
}

// Own component of a b; ao is the partner's part of a. For a and b in
// [0, 2p) part 0 reduces to (-p, 2p) and part 1 to [0, 3p); part 1 takes
// -p R into its sum, which puts it in [-p, 2p) as well, so one conditional
// addition of 2p brings either to [0, 2p). RAW (operands of magnitude below
// 2p): both parts lie in (-2p, 3p), corrected in both directions.
template <bool CHAIN = false, bool RAW = false>
__device__ __forceinline__ void fp2_lane_mul(Fp& r, const Fp& a, const Fp& ao, const Fp& b) {
    int32_t y[NL], w[NL];
    lane_operands(y, w, b, false);
    const Fp* const x[2] = {&a, &ao};
    const int32_t(*const yy[2])[NL] = {&y, &w};
    uint32_t t[NL];
    int64_t top;
    if constexpr (RAW) {
        redc_signed<2, CHAIN>(t, top, x, yy);
        fold_est<-1, 1>(r, t, (int32_t)top);
    } else {
        redc_signed<2, CHAIN, true>(t, top, x, yy, lane_part() != 0 ? -1 : 0);
        redc_fold(r, t, top);
    }
}

// Own component of a b - c d with one reduction. With every operand in
// [0, 2p) both parts lie in (-8p^2, 8p^2) before and in (-2p, 3p) after the
// reduction. `wide`: a and b raw (magnitude below 2p), c and d in [0, 2p);
// part 0 (a0 b0 - a1 b1 - c0 d0 + c1 d1) lies in (-12p^2, 12p^2) and part 1
// (a1 b0 + a0 b1 - c1 d0 - c0 d1) in (-16p^2, 8p^2), so with p < 0.238 R
// the reduced values lie in (-2.86p, 3.86p) and (-3.81p, 2.91p), within the
// (-4p, 4p) that redc_fold_wide takes.
template <bool CHAIN = false>
__device__ __forceinline__ void fp2_lane_mul_sub(Fp& r, const Fp& a, const Fp& ao, const Fp& b, const Fp& c,
                                                 const Fp& co, const Fp& d, bool wide = false) {
    int32_t y0[NL], w0[NL], y1[NL], w1[NL];
    lane_operands(y0, w0, b, false);
    lane_operands(y1, w1, d, true);
    const Fp* const x[4] = {&a, &ao, &c, &co};
    const int32_t(*const yy[4])[NL] = {&y0, &w0, &y1, &w1};
    uint32_t t[NL];
    int64_t top;
    redc_signed<4, CHAIN>(t, top, x, yy);
    if (wide)
        redc_fold_wide(r, t, top);
    else
        fold_est<-1, 1>(r, t, (int32_t)top);
}

// Own component of a^2 for a raw a (limbs of magnitude below 2^28, value of
// magnitude below 2p): part 0 (a0 + a1)(a0 - a1) = a0^2 - a1^2 in
// (-4p^2, 4p^2), part 1 (a1 + a1) a0 in (-8p^2, 8p^2), one signed product per
// lane (limb products below 2^58, a column below 2^63). The reduced value
// lies in (-2p, 3p); returned as t / top before the fold.
template <bool CHAIN = false>
__device__ __forceinline__ void fp2_lane_sqr_cols(uint32_t (&t)[NL], int64_t& top, const Fp& a) {
    Fp x;
    // part 0: x = a0 + a1, y = a0 - a1; part 1: x = a1 + a1, y = a0
    const int32_t m = lane_part() != 0 ? 0 : -1;
    int32_t y[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
        const int32_t a1 = (int32_t)pair_part1(a.v[i]);
        x.v[i] = (uint32_t)((int32_t)a.v[i] + a1);
        y[i] = (int32_t)pair_part0(a.v[i]) - (a1 & m);
    }
    const Fp* const xs[1] = {&x};
    const int32_t(*const ys[1])[NL] = {&y};
    redc_signed<1, CHAIN>(t, top, xs, ys);
}

// CHAIN selects the single-chain column sums of redc_signed (the
// accumulation kernel's policy, Fp2LaneAccOps).
template <bool CHAIN>
struct Fp2LaneOpsT {
    using F = Fp;
    static constexpr int WORDS = NL;
    static constexpr bool RAW = true;  // has sqr_sub2 / sqr_sub2z
    __device__ static __forceinline__ void add(F& r, const F& a, const F& b) { fp_add(r, a, b); }
    __device__ static __forceinline__ void sub(F& r, const F& a, const F& b) { fp_sub(r, a, b); }
    __device__ static __forceinline__ void dbl(F& r, const F& a) { fp_dbl(r, a); }
    __device__ static __forceinline__ void neg(F& r, const F& a) { fp_neg(r, a); }
    // own component of -a for a canonical a (the resident bases' coordinates)
    __device__ static __forceinline__ void neg_canon(F& r, const F& a) { fp_neg_canon(r, a); }
    __device__ static __forceinline__ void mul(F& r, const F& a, const F& b) {
        Fp ao;
        fp_pair_swap(ao, a);
        fp2_lane_mul<CHAIN>(r, a, ao, b);
    }
    // (a0 + a1 i)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 i: one plain product per
    // lane (part 0: (a0 + a1)(a0 - a1), part 1: (a1 + a1) a0), operands
    // picked per lane by selects so both lanes run one instruction stream.
    __device__ static __forceinline__ void sqr(F& r, const F& a) {
        Fp ao, x, y, t;
        fp_pair_swap(ao, a);
        const bool q = lane_part() != 0;
#pragma unroll
        for (int i = 0; i < NL; ++i) t.v[i] = q ? a.v[i] : ao.v[i];
        fp_add(x, a, t);
        fp_sub(y, a, ao);
#pragma unroll
        for (int i = 0; i < NL; ++i) y.v[i] = q ? ao.v[i] : y.v[i];
        fp_mul(r, x, y);
    }
    __device__ static __forceinline__ void mul_sub(F& r, const F& a, const F& b, const F& c, const F& d) {
        Fp ao, co;
        fp_pair_swap(ao, a);
        fp_pair_swap(co, c);
        fp2_lane_mul_sub<CHAIN>(r, a, ao, b, c, co, d);
    }
    // Raw differences, as FpOps has them: a difference that only feeds
    // products stays limb by limb (signed limbs, value of magnitude below
    // 2p), and the lane products take signed operands anyway. mul_raw takes
    // raw operands (each part's sum below 8p^2 in magnitude, two
    // corrections); mul_sub_raw (raw a, b) and the fused X3 fold wider ranges.
    __device__ static __forceinline__ void sub_raw(F& r, const F& a, const F& b) { fp_sub_raw(r, a, b); }
    __device__ static __forceinline__ void sqr_raw(F& r, const F& a) {
        uint32_t t[NL];
        int64_t top;
        fp2_lane_sqr_cols<CHAIN>(t, top, a);
        fold_est<-1, 1>(r, t, (int32_t)top);
    }
    __device__ static __forceinline__ void mul_raw(F& r, const F& a, const F& b) {
        Fp ao;
        fp_pair_swap(ao, a);
        fp2_lane_mul<CHAIN, true>(r, a, ao, b);
    }
    __device__ static __forceinline__ void mul_sub_raw(F& r, const F& a, const F& b, const F& c, const F& d) {
        Fp ao, co;
        fp_pair_swap(ao, a);
        fp_pair_swap(co, c);
        fp2_lane_mul_sub<CHAIN>(r, a, ao, b, c, co, d, true);
    }
    // Own component of a^2 - b - 2c (a raw, b and c in [0, 2p)): the square
    // reduces to (-2p, 3p); 8p - b - 2c (limbs 8 p_i - b_i - 2 c_i, in
    // (-3 * 2^28, 2^31)) is added with the carries, so the sum lies in
    // (0, 11p), and one quotient-estimate pass (fold_est<0, 5>) brings it to
    // [0, 2p).
    __device__ static __forceinline__ void sqr_sub2(F& r, const F& a, const F& b, const F& c) {
        uint32_t t[NL];
        int64_t top;
        fp2_lane_sqr_cols<CHAIN>(t, top, a);
        int64_t acc = 0;
#pragma unroll
        for (int i = 0; i < NL - 1; ++i) {
            acc += (int64_t)t[i] + ((int64_t)(int32_t)(8u * PL[i]) - (int32_t)b.v[i] - 2 * (int64_t)(int32_t)c.v[i]);
            t[i] = (uint32_t)acc & MASK28;
            acc >>= 28;
        }
        acc += top + ((int64_t)(8u * PL[NL - 1]) - (int32_t)b.v[NL - 1] - 2 * (int64_t)(int32_t)c.v[NL - 1]);
        fold_est<0, 5>(r, t, (int32_t)acc);  // in (0, 11p): non-negative, below 2^31
    }
    __device__ static __forceinline__ void sqr_sub2z(F& r, const F& a, const F& c) {
        Fp z;
        fp_set_zero(z);
        sqr_sub2(r, a, z, c);
    }
    __device__ static __forceinline__ bool is_zero(const F& a) {
        const uint32_t z = fp_is_zero(a) ? 1u : 0u;
        return (z & pair_swap(z)) != 0;
    }
    __device__ static __forceinline__ bool is_raw_zero(const F& a) {
        const uint32_t z = fp_is_raw_zero(a) ? 1u : 0u;
        return (z & pair_swap(z)) != 0;
    }
    __device__ static __forceinline__ void canon(F& r, const F& a) { fp_canon(r, a); }
    // (a0 - a1 i) / (a0^2 + a1^2): both lanes form the norm and invert it
    __device__ static __forceinline__ void inv(F& r, const F& a) {
        Fp s, so, n;
        fp_sqr(s, a);
        fp_pair_swap(so, s);
        fp_add(n, s, so);
        fp_inv(n, n);
        fp_mul(r, a, n);
        if (lane_part()) fp_neg(r, r);
    }
    __device__ static __forceinline__ void zero(F& r) { fp_set_zero(r); }
    __device__ static __forceinline__ void one(F& r) {
        if (lane_part())
            fp_set_zero(r);
        else
            fp_set_one(r);
    }
};
using Fp2LaneOps = Fp2LaneOpsT<false>;
using Fp2LaneAccOps = Fp2LaneOpsT<true>;


// G1 policy of the depth-bound tail kernels (bucket reduction, window sums):
// lanes 2k and 2k+1 both hold the full operands of one point operation and
// split its independent field products between them (part 0 and part 1 of
// each step), swapping results with the same DPP quad permutation as
// Fp2LaneOps. The field operations are FpOps'; the point formulas are
// specialised for this policy (xyzz_add / xyzz_dbl), and every branch depends
// on values both lanes hold, so control flow is uniform over each pair.
struct FpPairOps : FpOps {};

// r = c ? a : b, limb by limb.
__device__ __forceinline__ void fp_sel(Fp& r, bool c, const Fp& a, const Fp& b) {
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = c ? a.v[i] : b.v[i];
}

// After a paired step in which part 0 computed r0 and part 1 computed r1
// (each lane holding its own result in `mine`), gives both lanes both.
__device__ __forceinline__ void pair_results(Fp& r0, Fp& r1, const Fp& mine) {
    Fp other;
    fp_pair_swap(other, mine);
    const bool q = lane_part() != 0;
    fp_sel(r0, q, other, mine);
    fp_sel(r1, q, mine, other);
}

// G2 policy of the tail kernels: lanes 4k..4k+3 form two Fp2LaneOps pairs
// (lanes 4k, 4k+1 and 4k+2, 4k+3) that hold the same operands and compute
// the two independent Fp2 products of each step of a point operation, one
// per pair (half 0 and half 1), exchanging results with the DPP quad
// permutation [2, 3, 0, 1].
struct Fp2QuadOps : Fp2LaneOps {};

__device__ __forceinline__ uint32_t quad_half() { return (threadIdx.x >> 1) & 1u; }

// r = the value of a in the other pair of the quad (quad_perm [2, 3, 0, 1]).
__device__ __forceinline__ void fp_half_swap(Fp& r, const Fp& a) {
#pragma unroll
    for (int i = 0; i < NL; ++i) r.v[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)a.v[i], 0x4E, 0xF, 0xF, false);
}

// After a step in which half 0 computed r0 and half 1 computed r1, gives
// both halves both (each lane its own Fp2 component).
__device__ __forceinline__ void quad_results(Fp& r0, Fp& r1, const Fp& mine) {
    Fp other;
    fp_half_swap(other, mine);
    const bool h = quad_half() != 0;
    fp_sel(r0, h, other, mine);
    fp_sel(r1, h, mine, other);
}

}  // namespace gm
