// Engine<O>: orchestration of the per-stage launchers (no kernels here).
#include "engine_impl.hpp"

namespace gm {

template <class O>
hipError_t Engine<O>::run(hipStream_t s, const Plan& p, const Aff<O>* d_bases, const uint64_t* d_scalars,
                          uint8_t* ws, Prof* prof) {
    auto mark = [&](int st) -> hipError_t {
        return (prof && prof->on) ? hipEventRecord(prof->ev[st], s) : hipSuccess;
    };
    GM_TRY(mark(ST_START));
    GM_TRY(launch_sort(s, p, d_scalars, ws, prof));
    GM_TRY(launch_accumulate<O>(s, p, d_bases, ws));
    GM_TRY(mark(ST_ACCUM));
    GM_TRY(launch_fixup<O>(s, p, ws));
    GM_TRY(mark(ST_FIXUP));
    GM_TRY(launch_reduce<O>(s, p, ws));
    GM_TRY(mark(ST_REDUCE));
    GM_TRY(launch_window_sum<O>(s, p, ws));
    GM_TRY(mark(ST_WSUM));
    return hipSuccess;
}

template <class O>
hipError_t Engine<O>::device_combine(hipStream_t s, const Plan& p, uint8_t* ws) {
    return launch_horner<O>(s, reinterpret_cast<const MsmDesc*>(ws + p.o_desc), p.count,
                            reinterpret_cast<const Xyzz<O>*>(ws + p.o_wsum), reinterpret_cast<Aff<O>*>(ws + p.o_out));
}

template <class O>
hipError_t Engine<O>::import_aff(hipStream_t s, Aff<O>* out, const void* d_abi, int n, bool to_mont) {
    return launch_import<O>(s, out, d_abi, n, to_mont);
}

template <class O>
hipError_t Engine<O>::export_aff(hipStream_t s, void* d_abi, const Aff<O>* in, int n) {
    return launch_export<O>(s, d_abi, in, n);
}

template <class O>
hipError_t Engine<O>::gen_bases(hipStream_t s, Aff<O>* out, const uint64_t* d_k, int n, const Aff<O>& g) {
    return launch_gen_bases<O>(s, out, d_k, n, g);
}

template <class O>
hipError_t Engine<O>::aff_to_abi_proj(hipStream_t s, void* d_abi_proj, const Aff<O>* in) {
    return launch_aff_to_abi_proj<O>(s, d_abi_proj, in, 1);
}

template struct Engine<FpOps>;
template struct Engine<Fp2Ops>;

}  // namespace gm
