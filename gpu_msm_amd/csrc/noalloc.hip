// An MSM entry point that allocates nothing: the device-pointer MSM of
// point_msm_async<PointType> (affine result, multi-limb scalars) with its
// workspace in a fixed, file-scope device array instead of caller or
// stream-ordered memory, and the window combine on the device, so no pinned
// host slot either. For callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// In host code the name `scratch` is only a stand-in, not a device address:
// the array's device address is looked up once per device with
// hipGetSymbolAddress (that device current) and kept per device, and kernels
// get that address. Every call on one GPU uses the same array, which
// keeps the previous call's contents (the pipeline clears what must start at
// zero with fill kernels), so calls that overlap on one GPU race and must be
// serialised by the caller.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>
#include <type_traits>

#include "../../include/gpu_msm_ext.h"
#include "surface.hpp"

namespace {

// 144 MiB: room for G1 MSMs up to 2^16 points and G2 up to 2^15 with 5-limb
// scalars (the limit is the plan's exact workspace plus the imported bases).
constexpr size_t SCRATCH_BYTES = (size_t)144 << 20;
static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t* g_ptr[MAX_DEVICES];

// The device address of `scratch` on `gpu` (current device), looked up on
// the first call for that device.
hipError_t scratch_ptr(int gpu, uint8_t** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[gpu]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_ptr[gpu] = static_cast<uint8_t*>(p);
    }
    *out = g_ptr[gpu];
    return hipSuccess;
}

template <class O>
int msm_noalloc(hipStream_t s, uint8_t* ws_base, void* d_result, const void* d_points, const uint64_t* d_scalars,
                int limbs, int n) {
    using gm::Aff;
    if (n == 0)  // the result is O, as for the reference's n == 0 (src/curve.cu:2766-2769)
        return gm::launch_point_op<O>(s, gm::PO_INF, d_result, nullptr, nullptr, nullptr, 0, 0) == hipSuccess ? 0 : -4;
    gm::Plan p;
    gm::MsmDesc d;
    gm::MsmSpec spec = {n, 0, 0, 0, 0};
    const int curve = std::is_same<O, gm::FpOps>::value ? gm::CURVE_G1 : gm::CURVE_G2;
    if (!gm::make_plan(p, &d, curve, limbs, 64 * limbs, &spec, 1, true)) return -1;
    const size_t b_aff = gm::al_size((size_t)n * sizeof(Aff<O>));
    if (b_aff + p.total > SCRATCH_BYTES) return -6;
    Aff<O>* d_aff = reinterpret_cast<Aff<O>*>(ws_base);
    uint8_t* ws = ws_base + b_aff;
    if (gm::launch_put_desc(s, reinterpret_cast<gm::MsmDesc*>(ws + p.o_desc), d) != hipSuccess) return -4;
    if (gm::Engine<O>::import_aff(s, d_aff, d_points, n, false) != hipSuccess) return -4;
    if (gm::Engine<O>::run(s, p, d_aff, d_scalars, ws, nullptr) != hipSuccess) return -4;
    if (gm::Engine<O>::device_combine(s, p, ws) != hipSuccess) return -4;
    if (gm::Engine<O>::export_aff(s, d_result, reinterpret_cast<const Aff<O>*>(ws + p.o_out), 1) != hipSuccess)
        return -4;
    return 0;
}

}  // namespace

int gm_point_msm_noalloc(void* stream, uint32_t gpu_index, int curve, void* d_result, const void* d_points,
                         const uint64_t* d_scalars, int scalar_limbs, int n) {
    if (gpu_index >= MAX_DEVICES || (curve != GM_CURVE_G1 && curve != GM_CURVE_G2) || !d_result || n < 0) return -1;
    if (n > 0 && (!d_points || !d_scalars || scalar_limbs < 1 || scalar_limbs > gm::MAX_SCALAR_LIMBS)) return -1;
    if (hipSetDevice((int)gpu_index) != hipSuccess) return -2;
    uint8_t* ws = nullptr;
    if (scratch_ptr((int)gpu_index, &ws) != hipSuccess) return -2;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return curve == GM_CURVE_G1
               ? msm_noalloc<gm::FpOps>(s, ws, d_result, d_points, d_scalars, scalar_limbs, n)
               : msm_noalloc<gm::Fp2Ops>(s, ws, d_result, d_points, d_scalars, scalar_limbs, n);
}
