// Point format and validation kernel instantiations (g1).
#include "engine_impl.hpp"
#include "format.hpp"
#include "format_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_fmt(hipStream_t s, int op, const void* in, void* out, int n, int arg, int32_t* status) {
    if (n <= 0) return hipSuccess;
    const uint64_t* a = reinterpret_cast<const uint64_t*>(in);
    uint64_t* b = reinterpret_cast<uint64_t*>(out);
    switch (op) {
        case FMT_OP_SERIALIZE:
            k_fmt_serialize<O><<<blocks_for(n, 64), 64, 0, s>>>(a, b, n, arg);
            break;
        case FMT_OP_DESERIALIZE_C:
        case FMT_OP_DESERIALIZE_U:
            k_fmt_deserialize<O><<<blocks_for(n, 64), 64, 0, s>>>(a, b, n, op == FMT_OP_DESERIALIZE_C, arg, status);
            break;
        case FMT_OP_CHECK:
            k_fmt_check<O><<<blocks_for(n, 64), 64, 0, s>>>(a, n, arg, status);
            break;
        case FMT_OP_CANONICAL:
            k_fmt_canonical<O><<<blocks_for(n, 128), 128, 0, s>>>(a, b, n, arg);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_fmt<FpOps>(hipStream_t, int, const void*, void*, int, int, int32_t*);

}  // namespace gm
