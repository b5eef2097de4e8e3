// Host-side interface of the point format / validation kernels.
#pragma once
#include "engine.hpp"

namespace gm {

enum FmtOp : int {
    FMT_OP_SERIALIZE = 0,      // arg: compressed
    FMT_OP_DESERIALIZE_C = 1,  // arg: checks (FMT_CHECK_*)
    FMT_OP_DESERIALIZE_U = 2,  // arg: checks
    FMT_OP_CHECK = 3,          // arg: checks
    FMT_OP_CANONICAL = 4,      // arg: 1 Montgomery -> standard form, 0 the reverse
};

template <class O>
hipError_t launch_fmt(hipStream_t s, int op, const void* in, void* out, int n, int arg, int32_t* status);

}  // namespace gm
