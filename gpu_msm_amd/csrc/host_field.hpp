// Host-side BLS12-446 field and group arithmetic (x86-64, 7 x 64-bit limbs).
//
// Used by the library's host code: the final window combine of a single MSM
// (a ~300-doubling serial chain that a CPU core runs far faster than one GPU
// lane), normalisation to affine, and the ABI helpers of reference
// rust-api/src/c_wrapper.cu:12-59,261-288 (affine<->projective, Montgomery
// conversion). Values are canonical (< p), Montgomery form with R = 2^448, the
// representation of reference src/fp.cu.
#pragma once
#include <stdint.h>
#include <string.h>
#if defined(__x86_64__)
#include <x86intrin.h>
#endif

namespace gmh {

typedef unsigned __int128 u128;

struct Fp {
    uint64_t l[7];
};
struct Fp2 {
    Fp c0, c1;
};

static const Fp P = {{0x311c0026aab0aaabULL, 0x56ee4528c573b5ccULL, 0x824e6dc3e23acdeeULL, 0x0f75a64bbac71602ULL,
                      0x0095a4b78a02fe32ULL, 0x200fc34965aad640ULL, 0x3cdee0fb28c5e535ULL}};
static const Fp R2 = {{0x2AFF01DDDC752B45ULL, 0x92C772A7421CCF5BULL, 0x140EEF29C347DAD6ULL, 0xF5A1400C22EA595EULL,
                       0x99D91C9FEC145218ULL, 0x3BB6537F90143D4BULL, 0x3627854C9BE7974FULL}};
static const Fp ONE = {{0x3b8fff65553d5554ULL, 0xa446eb5cea3128cfULL, 0xf6c648f07714c846ULL, 0xc22966d114e3a7f5ULL,
                        0xfda96d21d7f40737ULL, 0x7fc0f2da6954a6ffULL, 0x0c847c135ce86b2bULL}};
static const uint64_t PINV = 0xcd63fd900035fffdULL;

inline bool geq_p(const uint64_t* a) {
    for (int i = 6; i >= 0; --i) {
        if (a[i] != P.l[i]) return a[i] > P.l[i];
    }
    return true;
}
inline void sub_p(uint64_t* a) {
    uint64_t br = 0;
    for (int i = 0; i < 7; ++i) {
        u128 d = (u128)a[i] - P.l[i] - br;
        a[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) & 1;
    }
}
inline bool is_zero(const Fp& a) {
    uint64_t o = 0;
    for (int i = 0; i < 7; ++i) o |= a.l[i];
    return o == 0;
}
inline bool eq(const Fp& a, const Fp& b) { return memcmp(a.l, b.l, sizeof a.l) == 0; }

// Carry and borrow steps of the word chains (adc / sbb on x86-64).
inline unsigned char addc(unsigned char c, uint64_t a, uint64_t b, uint64_t& o) {
#if defined(__x86_64__)
    unsigned long long t;
    c = _addcarry_u64(c, a, b, &t);
    o = t;
    return c;
#else
    const u128 s = (u128)a + b + c;
    o = (uint64_t)s;
    return (unsigned char)(s >> 64);
#endif
}
inline unsigned char subb(unsigned char br, uint64_t a, uint64_t b, uint64_t& o) {
#if defined(__x86_64__)
    unsigned long long t;
    br = _subborrow_u64(br, a, b, &t);
    o = t;
    return br;
#else
    const u128 d = (u128)a - b - br;
    o = (uint64_t)d;
    return (unsigned char)((d >> 64) & 1);
#endif
}

// t - p unless that borrows (t below 2p), masked rather than branched.
inline void sub_p_once(uint64_t* r, const uint64_t* t) {
    uint64_t d[7];
    unsigned char br = 0;
    for (int i = 0; i < 7; ++i) br = subb(br, t[i], P.l[i], d[i]);
    const uint64_t keep = 0 - (uint64_t)br;
    for (int i = 0; i < 7; ++i) r[i] = (t[i] & keep) | (d[i] & ~keep);
}

// a + b, reduced once: the sum minus p is kept unless that borrows (and no
// carry left the seventh word). No branch on the values.
inline void add(Fp& r, const Fp& a, const Fp& b) {
    uint64_t t[7], d[7];
    unsigned char c = 0, br = 0;
    for (int i = 0; i < 7; ++i) c = addc(c, a.l[i], b.l[i], t[i]);
    for (int i = 0; i < 7; ++i) br = subb(br, t[i], P.l[i], d[i]);
    const uint64_t keep = 0 - (uint64_t)(br & (c ^ 1));  // all ones: t is already below p
    for (int i = 0; i < 7; ++i) r.l[i] = (t[i] & keep) | (d[i] & ~keep);
}
// a - b, plus p where that borrows (masked, no branch).
inline void sub(Fp& r, const Fp& a, const Fp& b) {
    uint64_t t[7];
    unsigned char br = 0, c = 0;
    for (int i = 0; i < 7; ++i) br = subb(br, a.l[i], b.l[i], t[i]);
    const uint64_t m = 0 - (uint64_t)br;
    for (int i = 0; i < 7; ++i) c = addc(c, t[i], P.l[i] & m, r.l[i]);
}
// -a as p - a, and 0 for 0 (reference src/fp.cu:436-443).
inline void neg(Fp& r, const Fp& a) {
    if (is_zero(a)) {
        r = a;
        return;
    }
    sub(r, P, a);
}
// CIOS Montgomery product, canonical output, in x86-64 assembly (BMI2 mulx,
// ADX adcx/adox: two independent carry chains per row). p's top word is below
// 2^62, so the running value never needs the two extra words of textbook CIOS
// (the "no-carry" variant): with inputs below p each row's sum fits eight
// words (t0..t6 and A), every row's result stays below 2p in seven words, and
// one conditional subtraction ends it. Registers r8..r15 hold t and A; after
// each row t0 is zero, so the names rotate by one instead of shifting words:
// in row i, t_j is r[8 + (i + j) mod 8] and A is r[8 + (i + 7) mod 8].
// Compiled only for x86-64 and run only where the CPU reports BMI2 and ADX
// (mul() below); elsewhere the portable product serves.
#if defined(__x86_64__)
#define GMH_HAVE_ADX_PRODUCT 1
inline void mul_adx(Fp& r, const Fp& a, const Fp& b) {
    register uint64_t x8 asm("r8"), x9 asm("r9"), x10 asm("r10"), x11 asm("r11"), x12 asm("r12"),
        x13 asm("r13"), x14 asm("r14"), x15 asm("r15");
    asm volatile(
        "xorq %%r8, %%r8\n\t"
        "xorq %%r9, %%r9\n\t"
        "xorq %%r10, %%r10\n\t"
        "xorq %%r11, %%r11\n\t"
        "xorq %%r12, %%r12\n\t"
        "xorq %%r13, %%r13\n\t"
        "xorq %%r14, %%r14\n\t"
        "xorq %%r15, %%r15\n\t"
        "movq 0(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r15\n\t"
        "movq %%r8, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r15\n\t"
        "movq 8(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r8\n\t"
        "movq %%r9, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r8\n\t"
        "movq 16(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r9\n\t"
        "movq %%r10, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r9\n\t"
        "movq 24(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r10\n\t"
        "movq %%r11, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r10\n\t"
        "movq 32(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r11\n\t"
        "movq %%r12, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r11\n\t"
        "movq 40(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r12\n\t"
        "movq %%r13, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r13\n\t"
        "adcxq %%rbx, %%r14\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r12\n\t"
        "movq 48(%[b]), %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 8(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 16(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 24(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 32(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 40(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 48(%[a]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r13\n\t"
        "movq %%r14, %%rdx\n\t"
        "imulq %[pinv], %%rdx\n\t"
        "xorl %%eax, %%eax\n\t"
        "mulxq 0(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r14\n\t"
        "adcxq %%rbx, %%r15\n\t"
        "mulxq 8(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r15\n\t"
        "adcxq %%rbx, %%r8\n\t"
        "mulxq 16(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r8\n\t"
        "adcxq %%rbx, %%r9\n\t"
        "mulxq 24(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r9\n\t"
        "adcxq %%rbx, %%r10\n\t"
        "mulxq 32(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r10\n\t"
        "adcxq %%rbx, %%r11\n\t"
        "mulxq 40(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r11\n\t"
        "adcxq %%rbx, %%r12\n\t"
        "mulxq 48(%[p]), %%rax, %%rbx\n\t"
        "adoxq %%rax, %%r12\n\t"
        "adcxq %%rbx, %%r13\n\t"
        "movl $0, %%eax\n\t"
        "adoxq %%rax, %%r13\n\t"
        : "=&r"(x8), "=&r"(x9), "=&r"(x10), "=&r"(x11), "=&r"(x12), "=&r"(x13), "=&r"(x14), "=&r"(x15)
        : [a] "r"(a.l), [b] "r"(b.l), [p] "r"(P.l), [pinv] "m"(PINV)
        : "rax", "rbx", "rdx", "cc", "memory");
    const uint64_t t[7] = {x15, x8, x9, x10, x11, x12, x13};
    sub_p_once(r.l, t);
}
#endif

// The same CIOS product in portable C++ (128-bit intermediates), the same
// no-carry bound: rows stay below 2p in eight words.
inline void mul_portable(Fp& r, const Fp& a, const Fp& b) {
    uint64_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 7; ++i) {
        u128 c = 0;
        for (int j = 0; j < 7; ++j) {
            c += (u128)a.l[j] * b.l[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        t[7] = (uint64_t)c;
        const uint64_t m = t[0] * PINV;
        c = ((u128)m * P.l[0] + t[0]) >> 64;
        for (int j = 1; j < 7; ++j) {
            c += (u128)m * P.l[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        t[6] = t[7] + (uint64_t)c;
    }
    sub_p_once(r.l, t);
}

inline bool cpu_has_adx() {
#if defined(GMH_HAVE_ADX_PRODUCT)
    static const bool v = __builtin_cpu_supports("bmi2") && __builtin_cpu_supports("adx");
    return v;
#else
    return false;
#endif
}

inline void mul(Fp& r, const Fp& a, const Fp& b) {
#if defined(GMH_HAVE_ADX_PRODUCT)
    if (cpu_has_adx()) {
        mul_adx(r, a, b);
        return;
    }
#endif
    mul_portable(r, a, b);
}
inline void sqr(Fp& r, const Fp& a) { mul(r, a, a); }
inline void dbl(Fp& r, const Fp& a) { add(r, a, a); }
inline void to_mont(Fp& r, const Fp& a) { mul(r, a, R2); }
inline void from_mont(Fp& r, const Fp& a) {
    Fp one = {{1, 0, 0, 0, 0, 0, 0}};
    mul(r, a, one);
}
// Fermat inverse a^(p-2); 0 -> 0.
inline void inv(Fp& r, const Fp& a) {
    if (is_zero(a)) {
        r = a;
        return;
    }
    Fp e = P;
    e.l[0] -= 2;
    Fp acc = ONE;
    for (int i = 6; i >= 0; --i)
        for (int b = 63; b >= 0; --b) {
            sqr(acc, acc);
            if ((e.l[i] >> b) & 1) mul(acc, acc, a);
        }
    r = acc;
}
// Reduce an arbitrary 448-bit value mod p (at most four subtractions since
// 2^448 < 5p).
inline void reduce(Fp& r, const Fp& a) {
    r = a;
    while (geq_p(r.l)) sub_p(r.l);
}

// ---------------------------------------------------------------- Fp2 (i^2 = -1)
inline void add(Fp2& r, const Fp2& a, const Fp2& b) {
    add(r.c0, a.c0, b.c0);
    add(r.c1, a.c1, b.c1);
}
inline void sub(Fp2& r, const Fp2& a, const Fp2& b) {
    sub(r.c0, a.c0, b.c0);
    sub(r.c1, a.c1, b.c1);
}
inline void neg(Fp2& r, const Fp2& a) {
    neg(r.c0, a.c0);
    neg(r.c1, a.c1);
}
inline void dbl(Fp2& r, const Fp2& a) { add(r, a, a); }
inline void mul(Fp2& r, const Fp2& a, const Fp2& b) {
    Fp t0, t1, t2, t3;
    mul(t0, a.c0, b.c0);
    mul(t1, a.c1, b.c1);
    add(t2, a.c0, a.c1);
    add(t3, b.c0, b.c1);
    mul(t2, t2, t3);
    sub(r.c0, t0, t1);
    sub(t2, t2, t0);
    sub(r.c1, t2, t1);
}
// (a0 + a1 i)^2 = (a0 + a1)(a0 - a1) + 2 a0 a1 i: two products
inline void sqr(Fp2& r, const Fp2& a) {
    Fp s, d, t;
    add(s, a.c0, a.c1);
    sub(d, a.c0, a.c1);
    mul(t, a.c0, a.c1);
    mul(r.c0, s, d);
    add(r.c1, t, t);
}
inline bool is_zero(const Fp2& a) { return is_zero(a.c0) && is_zero(a.c1); }
inline bool eq(const Fp2& a, const Fp2& b) { return eq(a.c0, b.c0) && eq(a.c1, b.c1); }
inline void inv(Fp2& r, const Fp2& a) {
    Fp t0, t1, n;
    mul(t0, a.c0, a.c0);
    mul(t1, a.c1, a.c1);
    add(n, t0, t1);
    inv(n, n);
    mul(r.c0, a.c0, n);
    neg(t0, a.c1);
    mul(r.c1, t0, n);
}

// ------------------------------------------------------------- XYZZ group law
template <class F>
struct Xyzz {
    F X, Y, ZZ, ZZZ;
};

template <class F>
inline void set_inf(Xyzz<F>& p) {
    memset(&p, 0, sizeof p);
}
template <class F>
inline bool is_inf(const Xyzz<F>& p) {
    return is_zero(p.ZZ);
}

template <class F>
inline void xdbl(Xyzz<F>& r, const Xyzz<F>& p) {
    if (is_inf(p)) {
        r = p;
        return;
    }
    F U, V, W, S, M, t;
    dbl(U, p.Y);
    sqr(V, U);
    mul(W, U, V);
    mul(S, p.X, V);
    sqr(t, p.X);
    dbl(M, t);
    add(M, M, t);
    mul(r.ZZ, V, p.ZZ);
    mul(r.ZZZ, W, p.ZZZ);
    mul(W, W, p.Y);
    sqr(t, M);
    sub(t, t, S);
    sub(r.X, t, S);
    sub(t, S, r.X);
    mul(t, M, t);
    sub(r.Y, t, W);
}

template <class F>
inline void xadd(Xyzz<F>& r, const Xyzz<F>& p, const Xyzz<F>& q) {
    if (is_inf(p)) {
        r = q;
        return;
    }
    if (is_inf(q)) {
        r = p;
        return;
    }
    F U1, S1, P_, R_, PP, PPP, Q, t;
    mul(U1, p.X, q.ZZ);
    mul(P_, q.X, p.ZZ);
    sub(P_, P_, U1);
    mul(S1, p.Y, q.ZZZ);
    mul(R_, q.Y, p.ZZZ);
    sub(R_, R_, S1);
    if (is_zero(P_)) {
        if (is_zero(R_))
            xdbl(r, p);
        else
            set_inf(r);
        return;
    }
    sqr(PP, P_);
    mul(PPP, P_, PP);
    mul(Q, U1, PP);
    mul(t, p.ZZ, q.ZZ);
    mul(r.ZZ, t, PP);
    mul(t, p.ZZZ, q.ZZZ);
    mul(r.ZZZ, t, PPP);
    mul(S1, S1, PPP);
    sqr(t, R_);
    sub(t, t, PPP);
    sub(t, t, Q);
    sub(r.X, t, Q);
    sub(t, Q, r.X);
    mul(t, R_, t);
    sub(r.Y, t, S1);
}

// (x, y) = (X/ZZ, Y/ZZZ); false for infinity.
template <class F>
inline bool to_affine(F& x, F& y, const Xyzz<F>& p) {
    if (is_inf(p)) return false;
    F t, iv;
    mul(t, p.ZZ, p.ZZZ);
    inv(iv, t);
    mul(t, iv, p.ZZZ);
    mul(x, p.X, t);
    mul(t, iv, p.ZZ);
    mul(y, p.Y, t);
    return true;
}

// ------------------------------------------------- Jacobian group law (a = 0)
// x = X / Z^2, y = Y / Z^3, Z = 0 for infinity. The window combine runs its
// ~300 doublings here: a Jacobian doubling is 1M + 5S against 6M + 3S in XYZZ.
template <class F>
struct Jac {
    F X, Y, Z;
};

// XYZZ -> Jacobian without an inversion: ZZ^3 = ZZZ^2 makes
// (X ZZ^2, Y ZZ^3, ZZZ) a Jacobian representation of the same point.
template <class F>
inline void from_xyzz(Jac<F>& r, const Xyzz<F>& p) {
    if (is_inf(p)) {
        memset(&r, 0, sizeof r);
        return;
    }
    F z2, z3;
    sqr(z2, p.ZZ);
    mul(z3, z2, p.ZZ);
    mul(r.X, p.X, z2);
    mul(r.Y, p.Y, z3);
    r.Z = p.ZZZ;
}

// dbl-2009-l; Z = 0 or Y = 0 gives Z3 = 0 (infinity). r may alias p.
template <class F>
inline void jdbl(Jac<F>& r, const Jac<F>& p) {
    F A, B, C, D, E, Fv, t;
    sqr(A, p.X);
    sqr(B, p.Y);
    sqr(C, B);
    add(t, p.X, B);
    sqr(t, t);
    sub(t, t, A);
    sub(t, t, C);
    dbl(D, t);
    dbl(E, A);
    add(E, E, A);
    sqr(Fv, E);
    mul(t, p.Y, p.Z);
    dbl(r.Z, t);
    dbl(t, D);
    sub(r.X, Fv, t);
    sub(t, D, r.X);
    mul(t, E, t);
    dbl(C, C);
    dbl(C, C);
    dbl(C, C);
    sub(r.Y, t, C);
}

// add-2007-bl with the equal and opposite cases. r may alias p or q.
template <class F>
inline void jadd(Jac<F>& r, const Jac<F>& p, const Jac<F>& q) {
    if (is_zero(p.Z)) {
        r = q;
        return;
    }
    if (is_zero(q.Z)) {
        r = p;
        return;
    }
    F z1z1, z2z2, u1, u2, s1, s2, h, i, j, rr, v, t;
    sqr(z1z1, p.Z);
    sqr(z2z2, q.Z);
    mul(u1, p.X, z2z2);
    mul(u2, q.X, z1z1);
    mul(s1, p.Y, q.Z);
    mul(s1, s1, z2z2);
    mul(s2, q.Y, p.Z);
    mul(s2, s2, z1z1);
    sub(h, u2, u1);
    sub(rr, s2, s1);
    if (is_zero(h)) {
        if (is_zero(rr))
            jdbl(r, p);
        else
            memset(&r, 0, sizeof r);
        return;
    }
    dbl(rr, rr);
    dbl(i, h);
    sqr(i, i);
    mul(j, h, i);
    mul(v, u1, i);
    add(t, p.Z, q.Z);
    sqr(t, t);
    sub(t, t, z1z1);
    sub(t, t, z2z2);
    mul(r.Z, t, h);
    sqr(t, rr);
    sub(t, t, j);
    sub(t, t, v);
    sub(r.X, t, v);
    sub(t, v, r.X);
    mul(t, rr, t);
    mul(s1, s1, j);
    dbl(s1, s1);
    sub(r.Y, t, s1);
}

// (x, y) = (X / Z^2, Y / Z^3); false for infinity.
template <class F>
inline bool to_affine(F& x, F& y, const Jac<F>& p) {
    if (is_zero(p.Z)) return false;
    F iv, i2, i3;
    inv(iv, p.Z);
    sqr(i2, iv);
    mul(i3, i2, iv);
    mul(x, p.X, i2);
    mul(y, p.Y, i3);
    return true;
}

// ------------------------------------------------- 28-bit device limb packing
// Device elements are 16 x 28-bit limbs of the same 448-bit integer.
inline void from_limbs28(Fp& r, const uint32_t* v) {
    memset(r.l, 0, sizeof r.l);
    for (int i = 0; i < 16; ++i) {
        const int bit = 28 * i;
        const uint64_t x = v[i] & 0xFFFFFFFu;
        r.l[bit / 64] |= x << (bit % 64);
        if (bit % 64 > 36) r.l[bit / 64 + 1] |= x >> (64 - bit % 64);
    }
}
inline void to_limbs28(uint32_t* v, const Fp& a) {
    for (int i = 0; i < 16; ++i) {
        const int bit = 28 * i;
        uint64_t x = a.l[bit / 64] >> (bit % 64);
        if (bit % 64 > 36 && bit / 64 + 1 < 7) x |= a.l[bit / 64 + 1] << (64 - bit % 64);
        v[i] = (uint32_t)x & 0xFFFFFFFu;
    }
}

}  // namespace gmh
