// Shared helpers of the per-stage translation units.
#pragma once
#include "engine.hpp"

namespace gm {

#define GM_TRY(x)                        \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

static inline int blocks_for(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace gm
