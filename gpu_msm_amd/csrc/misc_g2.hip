// Format conversion, base generation and device combine kernels (g2).
#include "engine_impl.hpp"
#include "msm_kernels.hpp"

namespace gm {

template <class O>
hipError_t launch_horner(hipStream_t s, const MsmDesc* desc, int count, const Xyzz<O>* wsum, Aff<O>* out) {
    if (count <= 0) return hipSuccess;
    k_horner<O><<<blocks_for(count, 64), 64, 0, s>>>(desc, count, wsum, out);
    return hipGetLastError();
}

template <class O>
hipError_t launch_import(hipStream_t s, Aff<O>* out, const void* d_abi, int n, bool to_mont) {
    if (n <= 0) return hipSuccess;
    k_import_aff<O><<<blocks_for(n, 128), 128, 0, s>>>(out, reinterpret_cast<const uint8_t*>(d_abi), n, to_mont ? 1 : 0);
    return hipGetLastError();
}

template <class O>
hipError_t launch_export(hipStream_t s, void* d_abi, const Aff<O>* in, int n) {
    if (n <= 0) return hipSuccess;
    k_export_aff<O><<<blocks_for(n, 128), 128, 0, s>>>(reinterpret_cast<uint8_t*>(d_abi), in, n);
    return hipGetLastError();
}

template <class O>
hipError_t launch_gen_bases(hipStream_t s, Aff<O>* out, const uint64_t* d_k, int n, const Aff<O>& g) {
    if (n <= 0) return hipSuccess;
    k_gen_bases<O><<<blocks_for(n, 64), 64, 0, s>>>(out, d_k, n, g);
    return hipGetLastError();
}

template <class O>
hipError_t launch_aff_to_abi_proj(hipStream_t s, void* d_abi_proj, const Aff<O>* in, int count) {
    k_aff_to_abi_proj<O><<<blocks_for(count, 64), 64, 0, s>>>(reinterpret_cast<uint8_t*>(d_abi_proj), in, count);
    return hipGetLastError();
}

template hipError_t launch_horner<Fp2Ops>(hipStream_t, const MsmDesc*, int, const Xyzz<Fp2Ops>*, Aff<Fp2Ops>*);
template hipError_t launch_import<Fp2Ops>(hipStream_t, Aff<Fp2Ops>*, const void*, int, bool);
template hipError_t launch_export<Fp2Ops>(hipStream_t, void*, const Aff<Fp2Ops>*, int);
template hipError_t launch_gen_bases<Fp2Ops>(hipStream_t, Aff<Fp2Ops>*, const uint64_t*, int, const Aff<Fp2Ops>&);
template hipError_t launch_aff_to_abi_proj<Fp2Ops>(hipStream_t, void*, const Aff<Fp2Ops>*, int);

}  // namespace gm
