// Stage 1 of the MSM: signed-digit recoding, bucket counts, prefix sum and
// scatter into bucket-contiguous order (curve independent).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "engine_impl.hpp"
#include "sort_kernels.hpp"

namespace gm {

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
    if (n == 0) return hipSuccess;
    k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_sort(hipStream_t s, const Plan& p, const uint64_t* d_scalars, uint8_t* ws, Prof* prof) {
    const MsmDesc* desc = reinterpret_cast<const MsmDesc*>(ws + p.o_desc);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(ws + p.o_cnt);
    uint32_t* off = reinterpret_cast<uint32_t*>(ws + p.o_off);
    uint32_t* rank = reinterpret_cast<uint32_t*>(ws + p.o_rank);
    uint64_t* sorted = reinterpret_cast<uint64_t*>(ws + p.o_sorted);
    uint32_t* err = reinterpret_cast<uint32_t*>(ws + p.o_err);
    const size_t nbk = p.bkt_total;
    auto mark = [&](int st) -> hipError_t {
        return (prof && prof->on) ? hipEventRecord(prof->ev[st], s) : hipSuccess;
    };
    const bool tiled = p.lds_B > 0;
    const size_t lds = (size_t)p.lds_B * 4;
    uint16_t* dig = reinterpret_cast<uint16_t*>(ws + p.o_dig);
    uint32_t* tbase = reinterpret_cast<uint32_t*>(ws + p.o_tbase);
    uint32_t* blkcnt = reinterpret_cast<uint32_t*>(ws + p.o_blk);
    uint32_t* qbase = reinterpret_cast<uint32_t*>(ws + p.o_qbase);
    uint64_t* tmp = reinterpret_cast<uint64_t*>(ws + p.o_tmp);
    const uint32_t nblocks = (uint32_t)((nbk + SORT_BLOCK - 1) / SORT_BLOCK);
    // Counts (+1 sentinel so the scan also yields the total), the error words
    // (error flag, heavy-bucket count, entry count) and the block counts of
    // the two-pass placement are cleared by one fill kernel, not
    // hipMemsetAsync, so the sequence can be captured in a hipGraph and
    // replayed.
    {
        const size_t nblk = p.qmax ? nblocks : 0;
        const size_t longest = std::max<size_t>(nbk + 1, std::max<size_t>(4, nblk));
        k_zero3<<<(unsigned)((longest + 255) / 256), 256, 0, s>>>(cnt, nbk + 1, err, 4, blkcnt, nblk);
        GM_TRY(hipGetLastError());
    }
    if (tiled && lds > 65536) {
        // above the default dynamic LDS limit; once per process per kernel
        static const hipError_t attr = [] {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_tile_count),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 4 << (SORT_LDS_MAX_C - 1));
            if (e == hipSuccess)
                e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_tile_place),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 4 << (SORT_LDS_MAX_C - 1));
            return e;
        }();
        GM_TRY(attr);
    }
    if (p.n_total > 0) {
        if (tiled) {
            k_digits<<<blocks_for(p.n_total, 256), 256, 0, s>>>(desc, p.count, p.n_total, d_scalars, p.L, dig, err);
            GM_TRY(hipGetLastError());
            if (p.stile_total > 0) {
                k_tile_count<<<p.stile_total, 1024, lds, s>>>(desc, p.count, dig, cnt, tbase, blkcnt, qbase, p.qmax);
                GM_TRY(hipGetLastError());
            }
        } else {
            k_digits_hist<<<blocks_for(p.n_total, 256), 256, 0, s>>>(desc, p.count, p.n_total, d_scalars, p.L, cnt,
                                                                     rank, err);
            GM_TRY(hipGetLastError());
        }
    }
    GM_TRY(mark(ST_HIST));
    size_t cub_bytes = p.cub_bytes;
    GM_TRY(hipcub::DeviceScan::ExclusiveSum(ws + p.o_cub, cub_bytes, cnt, off, (int)(nbk + 1), s));
    GM_TRY(mark(ST_SCAN));
    if (p.n_total > 0) {
        if (tiled) {
            if (p.stile_total > 0 && p.qmax) {
                static const hipError_t attr1 = hipFuncSetAttribute(
                    reinterpret_cast<const void*>(k_tile_place_blocks), hipFuncAttributeMaxDynamicSharedMemorySize,
                    SORT_SUB * 10);
                GM_TRY(attr1);
                k_tile_place_blocks<<<p.stile_total, 1024, SORT_SUB * 10, s>>>(desc, p.count, dig, off, qbase, p.qmax,
                                                                              tmp);
                GM_TRY(hipGetLastError());
                static const hipError_t attr2 = hipFuncSetAttribute(reinterpret_cast<const void*>(k_block_place),
                                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                                    SORT_STAGE * 8);
                GM_TRY(attr2);
                k_block_place<<<nblocks, 1024, SORT_STAGE * 8, s>>>((uint32_t)nbk, off, tmp, sorted);
                GM_TRY(hipGetLastError());
            } else if (p.stile_total > 0) {
                k_tile_place<<<p.stile_total, 1024, lds, s>>>(desc, p.count, dig, off, tbase, sorted);
                GM_TRY(hipGetLastError());
            }
        } else {
            k_scatter<<<blocks_for(p.n_total, 256), 256, 0, s>>>(desc, p.count, p.n_total, d_scalars, p.L, off,
                                                                 rank, sorted);
            GM_TRY(hipGetLastError());
        }
    }
    GM_TRY(mark(ST_SCATTER));
    return hipSuccess;
}

}  // namespace gm
