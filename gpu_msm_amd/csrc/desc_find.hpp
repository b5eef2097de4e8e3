// Descriptor lookup shared by the batched kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "engine.hpp"

namespace gm {

// Index j of the MSM whose [field_j, field_{j+1}) range holds key (descriptor
// fields are ascending prefix sums), by binary search.
__device__ __forceinline__ int find_by(const uint32_t* __restrict__ base, int stride_words, int count, uint32_t key) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[(size_t)mid * stride_words] <= key)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

#define GM_FIND(d, count, field, key) \
    find_by(reinterpret_cast<const uint32_t*>(d) + offsetof(MsmDesc, field) / 4, sizeof(MsmDesc) / 4, count, key)

}  // namespace gm
