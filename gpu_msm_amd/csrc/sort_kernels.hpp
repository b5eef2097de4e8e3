// Signed-digit recoding and counting sort of (window, bucket) entries for a
// batch of MSMs (one descriptor per MSM).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "desc_find.hpp"

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

__global__ void __launch_bounds__(256) k_digits_hist(const MsmDesc* __restrict__ desc, int count, uint32_t n_total,
                                                     const uint64_t* __restrict__ scalars, int L,
                                                     uint32_t* __restrict__ cnt, uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ err) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_total) return;
    const MsmDesc& m = desc[GM_FIND(desc, count, pt_base, p)];
    const uint32_t i = p - m.pt_base;
    DigitIter it;
    it.load(scalars + (size_t)(m.scal_off + i) * L, L, (int)m.c);
    for (uint32_t w = 0; w < m.nw; ++w) {
        const int d = it.next();
        uint32_t r = 0xFFFFFFFFu;
        if (d != 0) {
            const uint32_t b = (uint32_t)(d < 0 ? -d : d) - 1;
            r = atomicAdd(&cnt[m.bkt_base + w * m.B + b], 1u);
        }
        rank[(size_t)m.rank_base + (size_t)w * m.n + i] = r;
    }
    if (!it.fits()) atomicOr(err, 1u);
}

__global__ void __launch_bounds__(256) k_scatter(const MsmDesc* __restrict__ desc, int count, uint32_t n_total,
                                                 const uint64_t* __restrict__ scalars, int L,
                                                 const uint32_t* __restrict__ off, const uint32_t* __restrict__ rank,
                                                 uint64_t* __restrict__ sorted) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_total) return;
    const MsmDesc& m = desc[GM_FIND(desc, count, pt_base, p)];
    const uint32_t i = p - m.pt_base;
    DigitIter it;
    it.load(scalars + (size_t)(m.scal_off + i) * L, L, (int)m.c);
    for (uint32_t w = 0; w < m.nw; ++w) {
        const int d = it.next();
        if (d != 0) {
            const uint32_t g = m.bkt_base + w * m.B + (uint32_t)(d < 0 ? -d : d) - 1;
            const uint32_t pos = off[g] + rank[(size_t)m.rank_base + (size_t)w * m.n + i];
            // (global bucket << 32) | base index in the pool | sign bit
            sorted[pos] = ((uint64_t)g << 32) | (m.base_off + i) | (d < 0 ? 0x80000000u : 0u);
        }
    }
}

// ------------------------------------------- tiled counting sort (c <= 16)
//
// k_digits      signed digits of every (window, point), coded in 16 bits:
//               (|d| - 1) | sign << 15, DIG_ZERO for d = 0 (|d| <= 2^15 and
//               |d| = 2^15 only occurs with a positive sign, so the code of
//               zero is free);
// k_tile_count  one workgroup per (MSM, window, tile of stile_pts points):
//               bucket counts in LDS, then one global atomic per non-empty
//               bucket of the tile, whose return value is the tile's base
//               inside the bucket;
// scan          exclusive prefix sum of the global counts;
// k_tile_place  the same tiles again: LDS counters start at
//               off[bucket] + tile base, and an LDS atomic per entry returns
//               its final position.
// Global atomics drop from one per entry to one per (tile, bucket).
//
// Windows of more than SORT_BLOCK buckets place their entries in two passes
// instead (plan qmax > 0): one LDS counter per bucket leaves every entry's
// 8-byte store scattered, which cost 0.25 of the 0.30 ms of k_tile_place at
// G1 2^20. k_tile_count then also gives each tile its slice of every block
// of SORT_BLOCK consecutive global buckets; k_tile_place_blocks moves the
// entries to their blocks, k_block_place sorts each block by bucket; both
// stage their entries in LDS and store whole runs (0.15 ms in all).
constexpr uint16_t DIG_ZERO = 0xFFFF;
// digits loaded ahead per thread in the tile kernels
constexpr int SORT_UNROLL = 8;
// SORT_QMAX: the most blocks of SORT_BLOCK buckets one window of
// 2^(SORT_LDS_MAX_C - 1) buckets touches (unaligned at both ends).
// entries of one block sorted in LDS by the second pass (128 KB)
constexpr uint32_t SORT_STAGE = 16384;
// points of a sort tile staged per step of the first pass (8 per thread of
// 1024; 80 KB of LDS with their block numbers)
constexpr uint32_t SORT_SUB = 8192;
constexpr uint32_t SORT_QMAX = (1u << (SORT_LDS_MAX_C - 1)) / SORT_BLOCK + 2;

__global__ void __launch_bounds__(256) k_digits(const MsmDesc* __restrict__ desc, int count, uint32_t n_total,
                                                const uint64_t* __restrict__ scalars, int L,
                                                uint16_t* __restrict__ dig, uint32_t* __restrict__ err) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_total) return;
    const MsmDesc& m = desc[GM_FIND(desc, count, pt_base, p)];
    const uint32_t i = p - m.pt_base;
    DigitIter it;
    it.load(scalars + (size_t)(m.scal_off + i) * L, L, (int)m.c);
    for (uint32_t w = 0; w < m.nw; ++w) {
        const int d = it.next();
        const uint16_t code =
            d == 0 ? DIG_ZERO : (uint16_t)(((uint32_t)(d < 0 ? -d : d) - 1) | (d < 0 ? 0x8000u : 0u));
        dig[(size_t)m.rank_base + (size_t)w * m.n + i] = code;
    }
    if (!it.fits()) atomicOr(err, 1u);
}

// (MSM, window, tile) of sort workgroup gt, and the tile's point range.
struct SortTile {
    uint32_t w, local, i0, i1;
};
__device__ __forceinline__ SortTile sort_tile(const MsmDesc& m, uint32_t gt) {
    const uint32_t per_w = (m.n + m.stile_pts - 1) / m.stile_pts;
    SortTile t;
    t.local = gt - m.stile_base;
    t.w = t.local / per_w;
    const uint32_t k = t.local - t.w * per_w;
    t.i0 = k * m.stile_pts;
    t.i1 = min(t.i0 + m.stile_pts, m.n);
    return t;
}

__global__ void __launch_bounds__(1024) k_tile_count(const MsmDesc* __restrict__ desc, int count,
                                                    const uint16_t* __restrict__ dig, uint32_t* __restrict__ cnt,
                                                    uint32_t* __restrict__ tbase, uint32_t* __restrict__ blkcnt,
                                                    uint32_t* __restrict__ qbase, uint32_t qmax) {
    extern __shared__ uint32_t lc[];
    __shared__ uint32_t bs[SORT_QMAX];
    const uint32_t gt = blockIdx.x;
    const MsmDesc& m = desc[GM_FIND(desc, count, stile_base, gt)];
    const SortTile t = sort_tile(m, gt);
    for (uint32_t b = threadIdx.x; b < m.B; b += blockDim.x) lc[b] = 0;
    __syncthreads();
    const uint16_t* dw = dig + (size_t)m.rank_base + (size_t)t.w * m.n;
    // SORT_UNROLL independent digit loads in flight per thread before their
    // use: one memory latency per SORT_UNROLL entries, not per entry
    for (uint32_t i = t.i0 + threadIdx.x; i < t.i1; i += SORT_UNROLL * blockDim.x) {
        uint32_t code[SORT_UNROLL];
#pragma unroll
        for (int u = 0; u < SORT_UNROLL; ++u) {
            const uint32_t j = i + u * blockDim.x;
            code[u] = j < t.i1 ? dw[j] : DIG_ZERO;
        }
#pragma unroll
        for (int u = 0; u < SORT_UNROLL; ++u)
            if (code[u] != DIG_ZERO) atomicAdd(&lc[code[u] & 0x7FFFu], 1u);
    }
    __syncthreads();
    uint32_t* gc = cnt + m.bkt_base + (size_t)t.w * m.B;
    if (qmax == 0) {
        uint32_t* tb = tbase + m.tb_base + (size_t)t.local * m.B;
        for (uint32_t b = threadIdx.x; b < m.B; b += blockDim.x) {
            const uint32_t c = lc[b];
            tb[b] = c ? atomicAdd(&gc[b], c) : 0u;
        }
        return;
    }
    // two-pass placement: bucket totals, and the tile's count and base in
    // every block of SORT_BLOCK global buckets its window touches
    const uint32_t g0 = m.bkt_base + t.w * m.B;
    const uint32_t q0 = g0 / SORT_BLOCK;
    const uint32_t nblk = (g0 + m.B - 1) / SORT_BLOCK - q0 + 1;  // <= B / SORT_BLOCK + 2 <= qmax
    for (uint32_t q = threadIdx.x; q < nblk; q += blockDim.x) bs[q] = 0;
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < m.B; b += blockDim.x) {
        const uint32_t c = lc[b];
        if (c) {
            atomicAdd(&gc[b], c);
            atomicAdd(&bs[(g0 + b) / SORT_BLOCK - q0], c);
        }
    }
    __syncthreads();
    uint32_t* qb = qbase + (size_t)gt * qmax;
    for (uint32_t q = threadIdx.x; q < nblk; q += blockDim.x) {
        const uint32_t c = bs[q];
        qb[q] = c ? atomicAdd(&blkcnt[q0 + q], c) : 0u;
    }
}

// Two-pass placement, pass 1: the tile's entries go to the range of their
// block of SORT_BLOCK global buckets (the tile's slice of it from
// k_tile_count). The tile is taken in steps of SORT_SUB points: count the
// step's entries per block, scan, stage them in LDS in block order, then
// store each block's run with consecutive lanes on consecutive slots, so the
// stores are whole lines rather than one scattered 8-byte write per entry.
__global__ void __launch_bounds__(1024) k_tile_place_blocks(const MsmDesc* __restrict__ desc, int count,
                                                           const uint16_t* __restrict__ dig,
                                                           const uint32_t* __restrict__ off,
                                                           const uint32_t* __restrict__ qbase, uint32_t qmax,
                                                           uint64_t* __restrict__ tmp) {
    extern __shared__ uint64_t stage[];  // SORT_SUB entries, then SORT_SUB block numbers (u16)
    uint16_t* sq = reinterpret_cast<uint16_t*>(stage + SORT_SUB);
    __shared__ uint32_t gpos[SORT_QMAX], lcnt[SORT_QMAX], lstart[SORT_QMAX], lrun[SORT_QMAX];
    __shared__ uint32_t step_total;
    constexpr int SCAN_PER = (SORT_QMAX + 63) / 64;
    const uint32_t gt = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const MsmDesc& m = desc[GM_FIND(desc, count, stile_base, gt)];
    const SortTile t = sort_tile(m, gt);
    const uint32_t g0 = m.bkt_base + t.w * m.B;
    const uint32_t q0 = g0 / SORT_BLOCK;
    const uint32_t nblk = (g0 + m.B - 1) / SORT_BLOCK - q0 + 1;
    const uint32_t* qb = qbase + (size_t)gt * qmax;
    for (uint32_t q = tid; q < nblk; q += blockDim.x) gpos[q] = off[(q0 + q) * SORT_BLOCK] + qb[q];
    const uint16_t* dw = dig + (size_t)m.rank_base + (size_t)t.w * m.n;
    constexpr int PER = SORT_SUB / 1024;  // entries per thread per step
    for (uint32_t base = t.i0; base < t.i1; base += SORT_SUB) {
        const uint32_t lim = min(base + SORT_SUB, t.i1);
        for (uint32_t q = tid; q < nblk; q += blockDim.x) lcnt[q] = 0;
        __syncthreads();
        uint32_t code[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const uint32_t j = base + tid + u * 1024;
            code[u] = j < lim ? dw[j] : DIG_ZERO;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (code[u] != DIG_ZERO) atomicAdd(&lcnt[(g0 + (code[u] & 0x7FFFu)) / SORT_BLOCK - q0], 1u);
        __syncthreads();
        // exclusive scan of the step's block counts by the first wave,
        // SCAN_PER counts per lane (nblk <= SORT_QMAX <= 64 * SCAN_PER)
        if (tid < 64) {
            uint32_t v[SCAN_PER], sum = 0;
#pragma unroll
            for (int k = 0; k < SCAN_PER; ++k) {
                const uint32_t q = tid * SCAN_PER + k;
                v[k] = q < nblk ? lcnt[q] : 0u;
                sum += v[k];
            }
            uint32_t inc = sum;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t o = __shfl_up(inc, d);
                if ((int)tid >= d) inc += o;
            }
            uint32_t ex = inc - sum;
#pragma unroll
            for (int k = 0; k < SCAN_PER; ++k) {
                const uint32_t q = tid * SCAN_PER + k;
                if (q < nblk) {
                    lstart[q] = ex;
                    lrun[q] = ex;
                }
                ex += v[k];
            }
            if (tid == 63) step_total = inc;
        }
        __syncthreads();
        const uint32_t total = step_total;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            if (code[u] == DIG_ZERO) continue;
            const uint32_t g = g0 + (code[u] & 0x7FFFu);
            const uint32_t q = g / SORT_BLOCK - q0;
            const uint32_t at = atomicAdd(&lrun[q], 1u);
            stage[at] = ((uint64_t)g << 32) | (m.base_off + base + tid + u * 1024) | ((code[u] & 0x8000u) << 16);
            sq[at] = (uint16_t)q;
        }
        __syncthreads();
        for (uint32_t k = tid; k < total; k += blockDim.x) {
            const uint32_t q = sq[k];
            tmp[gpos[q] + (k - lstart[q])] = stage[k];
        }
        __syncthreads();
        for (uint32_t q = tid; q < nblk; q += blockDim.x) gpos[q] += lcnt[q];
    }
}

// Pass 2: one workgroup per block of SORT_BLOCK global buckets sorts the
// block's entries (contiguous after pass 1) by bucket in LDS, one LDS counter
// per bucket, and writes the block back in its final order: consecutive
// lanes store consecutive entries, so the stores are whole lines. A block of
// more than SORT_STAGE entries (few, very full buckets) stores each entry at
// its final place directly.
__global__ void __launch_bounds__(1024) k_block_place(uint32_t bkt_total, const uint32_t* __restrict__ off,
                                                      const uint64_t* __restrict__ tmp,
                                                      uint64_t* __restrict__ sorted) {
    extern __shared__ uint64_t stage[];
    __shared__ uint32_t pos[SORT_BLOCK];
    const uint32_t gb = blockIdx.x * SORT_BLOCK;
    const uint32_t ge = min(gb + SORT_BLOCK, bkt_total);
    const uint32_t e0 = off[gb], e1 = off[ge];
    const bool staged = e1 - e0 <= SORT_STAGE;
    for (uint32_t b = threadIdx.x; b < ge - gb; b += blockDim.x) pos[b] = off[gb + b] - (staged ? e0 : 0u);
    __syncthreads();
    for (uint32_t j = e0 + threadIdx.x; j < e1; j += SORT_UNROLL * blockDim.x) {
        uint64_t e[SORT_UNROLL];
#pragma unroll
        for (int u = 0; u < SORT_UNROLL; ++u) {
            const uint32_t k = j + u * blockDim.x;
            e[u] = k < e1 ? tmp[k] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < SORT_UNROLL; ++u) {
            if (j + u * blockDim.x >= e1) continue;
            const uint32_t at = atomicAdd(&pos[(uint32_t)(e[u] >> 32) - gb], 1u);
            if (staged)
                stage[at] = e[u];
            else
                sorted[at] = e[u];
        }
    }
    if (!staged) return;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < e1 - e0; j += blockDim.x) sorted[e0 + j] = stage[j];
}

__global__ void __launch_bounds__(1024) k_tile_place(const MsmDesc* __restrict__ desc, int count,
                                                    const uint16_t* __restrict__ dig,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ tbase,
                                                    uint64_t* __restrict__ sorted) {
    extern __shared__ uint32_t pos[];
    const uint32_t gt = blockIdx.x;
    const MsmDesc& m = desc[GM_FIND(desc, count, stile_base, gt)];
    const SortTile t = sort_tile(m, gt);
    const uint32_t g0 = m.bkt_base + t.w * m.B;
    const uint32_t* tb = tbase + m.tb_base + (size_t)t.local * m.B;
    for (uint32_t b = threadIdx.x; b < m.B; b += blockDim.x) pos[b] = off[g0 + b] + tb[b];
    __syncthreads();
    const uint16_t* dw = dig + (size_t)m.rank_base + (size_t)t.w * m.n;
    for (uint32_t i = t.i0 + threadIdx.x; i < t.i1; i += SORT_UNROLL * blockDim.x) {
        uint32_t code[SORT_UNROLL];
#pragma unroll
        for (int u = 0; u < SORT_UNROLL; ++u) {
            const uint32_t j = i + u * blockDim.x;
            code[u] = j < t.i1 ? dw[j] : DIG_ZERO;
        }
#pragma unroll
        for (int u = 0; u < SORT_UNROLL; ++u) {
            if (code[u] == DIG_ZERO) continue;
            const uint32_t b = code[u] & 0x7FFFu;
            const uint32_t at = atomicAdd(&pos[b], 1u);
            // (global bucket << 32) | base index in the pool | sign bit
            sorted[at] = ((uint64_t)(g0 + b) << 32) | (m.base_off + i + u * blockDim.x) | ((code[u] & 0x8000u) << 16);
        }
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Zeroes three word ranges in one launch (the grid covers the longest).
__global__ void __launch_bounds__(256) k_zero3(uint32_t* __restrict__ a, size_t na, uint32_t* __restrict__ b,
                                               size_t nb, uint32_t* __restrict__ c, size_t nc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0;
    if (i < nb) b[i] = 0;
    if (i < nc) c[i] = 0;
}

}  // namespace gm
