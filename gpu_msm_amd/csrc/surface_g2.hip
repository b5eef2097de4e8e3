// Single-operation kernels (g2): point operations, Montgomery batch
// conversion and element-wise field operations.
#include "engine_impl.hpp"
#include "surface_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_point_op(hipStream_t s, int op, void* r, const void* a, const void* b, const uint64_t* d_sc,
                           uint64_t sc, int limbs) {
    k_point_op<O><<<1, 64, 0, s>>>(op, reinterpret_cast<uint64_t*>(r), reinterpret_cast<const uint64_t*>(a),
                                   reinterpret_cast<const uint64_t*>(b), d_sc, sc, limbs);
    return hipGetLastError();
}

template <class O>
hipError_t launch_point_mont_batch(hipStream_t s, void* pts, int n, bool to_mont) {
    if (n <= 0) return hipSuccess;
    k_point_mont_batch<O><<<blocks_for(n, 128), 128, 0, s>>>(reinterpret_cast<uint64_t*>(pts), n, to_mont ? 1 : 0);
    return hipGetLastError();
}

template <class O>
hipError_t launch_field_op(hipStream_t s, int op, void* c, const void* a, const void* b, int n, uint64_t arg,
                           int* iout) {
    if (n <= 0) return hipSuccess;
    k_field_op<O><<<blocks_for(n, 128), 128, 0, s>>>(op, reinterpret_cast<uint64_t*>(c),
                                                     reinterpret_cast<const uint64_t*>(a),
                                                     reinterpret_cast<const uint64_t*>(b), n, arg, iout);
    return hipGetLastError();
}

template hipError_t launch_point_op<Fp2Ops>(hipStream_t, int, void*, const void*, const void*, const uint64_t*, uint64_t,
                                         int);
template hipError_t launch_point_mont_batch<Fp2Ops>(hipStream_t, void*, int, bool);
template hipError_t launch_field_op<Fp2Ops>(hipStream_t, int, void*, const void*, const void*, int, uint64_t, int*);

}  // namespace gm
