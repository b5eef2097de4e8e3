// Signed-digit recoding and counting sort of (window, bucket) entries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gm {

// ---------------------------------------------------------- signed digits

// Signed recoding: d_w = bits_w + carry, folded into [-2^(c-1), 2^(c-1)].
// The scalar is consumed as a bit stream straight from global memory (no
// private array with a runtime index, which the compiler would move to LDS or
// scratch).
struct DigitIter {
    const uint64_t* p;
    uint64_t cur;      // unconsumed bits of the current limb (low `avail` bits)
    int avail, li, L, c;
    uint32_t carry;
    __device__ __forceinline__ void load(const uint64_t* ptr, int limbs, int cbits) {
        p = ptr;
        L = limbs;
        c = cbits;
        carry = 0;
        cur = L > 0 ? p[0] : 0;
        avail = 64;
        li = 1;
    }
    __device__ __forceinline__ uint32_t take() {
        const uint64_t mask = (1ull << c) - 1;
        uint64_t v;
        if (avail >= c) {
            v = cur & mask;
            cur = (c == 64) ? 0 : (cur >> c);
            avail -= c;
        } else {
            const uint64_t nxt = li < L ? p[li] : 0;
            ++li;
            const int need = c - avail;
            v = (cur | (avail ? (nxt << avail) : nxt)) & mask;
            cur = nxt >> need;
            avail = 64 - need;
        }
        return (uint32_t)v;
    }
    __device__ __forceinline__ int next() {
        const uint32_t half = 1u << (c - 1);
        const uint32_t raw = take() + carry;
        if (raw > half) {
            carry = 1;
            return (int)raw - (int)(half << 1);
        }
        carry = 0;
        return (int)raw;
    }
    // true if every bit after the consumed windows is zero and no carry remains.
    __device__ __forceinline__ bool fits() const {
        if (carry) return false;
        if (avail > 0 && cur != 0) return false;
        for (int j = li; j < L; ++j)
            if (p[j] != 0) return false;
        return true;
    }
};

__global__ void __launch_bounds__(256) k_digits_hist(const uint64_t* __restrict__ scalars, int n, int L, int c, int nw,
                              uint32_t* __restrict__ cnt, uint32_t* __restrict__ rank, uint32_t* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DigitIter it;
    it.load(scalars + (size_t)i * L, L, c);
    const int B = 1 << (c - 1);
    for (int w = 0; w < nw; ++w) {
        const int d = it.next();
        uint32_t r = 0xFFFFFFFFu;
        if (d != 0) {
            const int b = (d < 0 ? -d : d) - 1;
            r = atomicAdd(&cnt[(size_t)w * B + b], 1u);
        }
        rank[(size_t)w * n + i] = r;
    }
    if (!it.fits()) atomicOr(err, 1u);
}

__global__ void __launch_bounds__(256) k_scatter(const uint64_t* __restrict__ scalars, int n, int L, int c, int nw,
                          const uint32_t* __restrict__ off, const uint32_t* __restrict__ rank,
                          uint64_t* __restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DigitIter it;
    it.load(scalars + (size_t)i * L, L, c);
    const int B = 1 << (c - 1);
    for (int w = 0; w < nw; ++w) {
        const int d = it.next();
        if (d != 0) {
            const int b = (d < 0 ? -d : d) - 1;
            const uint32_t pos = off[(size_t)w * B + b] + rank[(size_t)w * n + i];
            // (global bucket << 32) | point index | sign bit
            sorted[pos] = ((uint64_t)((uint32_t)w * B + b) << 32) | (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
        }
    }
}

}  // namespace gm
