// Stage 1 of the MSM: signed-digit recoding, bucket counts, prefix sum and
// scatter into bucket-contiguous order (curve independent).
#include <hipcub/hipcub.hpp>

#include "engine_impl.hpp"
#include "sort_kernels.hpp"

namespace gm {

hipError_t launch_sort(hipStream_t s, const Plan& p, const uint64_t* d_scalars, uint8_t* ws, Prof* prof) {
    uint32_t* cnt = reinterpret_cast<uint32_t*>(ws + p.o_cnt);
    uint32_t* off = reinterpret_cast<uint32_t*>(ws + p.o_off);
    uint32_t* rank = reinterpret_cast<uint32_t*>(ws + p.o_rank);
    uint64_t* sorted = reinterpret_cast<uint64_t*>(ws + p.o_sorted);
    uint32_t* err = reinterpret_cast<uint32_t*>(ws + p.o_err);
    const size_t nbk = (size_t)p.nw * p.B;
    auto mark = [&](int st) -> hipError_t {
        return (prof && prof->on) ? hipEventRecord(prof->ev[st], s) : hipSuccess;
    };
    // counts (+1 sentinel so the scan also yields the total) and error flag
    GM_TRY(hipMemsetAsync(cnt, 0, (nbk + 1) * sizeof(uint32_t), s));
    GM_TRY(hipMemsetAsync(err, 0, sizeof(uint32_t), s));
    if (p.n > 0) {
        k_digits_hist<<<blocks_for(p.n, 256), 256, 0, s>>>(d_scalars, p.n, p.L, p.c, p.nw, cnt, rank, err);
        GM_TRY(hipGetLastError());
    }
    GM_TRY(mark(ST_HIST));
    size_t cub_bytes = p.cub_bytes;
    GM_TRY(hipcub::DeviceScan::ExclusiveSum(ws + p.o_cub, cub_bytes, cnt, off, (int)(nbk + 1), s));
    GM_TRY(mark(ST_SCAN));
    if (p.n > 0) {
        k_scatter<<<blocks_for(p.n, 256), 256, 0, s>>>(d_scalars, p.n, p.L, p.c, p.nw, off, rank, sorted);
        GM_TRY(hipGetLastError());
    }
    GM_TRY(mark(ST_SCATTER));
    return hipSuccess;
}

}  // namespace gm
