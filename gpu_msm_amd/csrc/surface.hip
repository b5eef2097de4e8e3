// The reference's C++ device-pointer surfaces on this library's kernels:
//   include/curve.h      point operations and the device-pointer MSM
//   include/fp.h         Fp stream operations
//   include/fp2.h        Fp2 stream operations
//   include/fp_kernels.h / fp2_kernels.h  element-wise arrays, host-pointer ops
// Argument checks and HIP failures abort() with a message naming the function,
// as the reference's PANIC_IF_FALSE / check_cuda_error do (reference
// include/device.h:11-39).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/curve.h"
#include "../../include/fp.h"
#include "../../include/fp2.h"
#include "../../include/fp2_kernels.h"
#include "../../include/fp_kernels.h"
#include "host_combine.hpp"
#include "surface.hpp"

using gm::Aff;
using gm::Fp2Ops;
using gm::FpOps;
using gm::Plan;
using gm::Xyzz;

namespace {

[[noreturn]] void panic(const char* fn, const char* what) {
    fprintf(stderr, "gpu_msm_amd: %s: %s\n", fn, what);
    fflush(stderr);
    abort();
}

void check(hipError_t e, const char* fn) {
    if (e == hipSuccess) return;
    fprintf(stderr, "gpu_msm_amd: %s: HIP error %d (%s)\n", fn, (int)e, hipGetErrorString(e));
    fflush(stderr);
    abort();
}

template <class T>
struct OpsOf;
template <>
struct OpsOf<G1Point> {
    using O = FpOps;
};
template <>
struct OpsOf<G2Point> {
    using O = Fp2Ops;
};
template <>
struct OpsOf<Fp> {
    using O = FpOps;
};
template <>
struct OpsOf<Fp2> {
    using O = Fp2Ops;
};

// ------------------------------------------------------------ points

template <class P>
void point_op(hipStream_t s, uint32_t gpu, int op, P* r, const P* a, const P* b, const uint64_t* d_sc, uint64_t sc,
              int limbs, bool sync, const char* fn) {
    check(hipSetDevice((int)gpu), fn);
    check(gm::launch_point_op<typename OpsOf<P>::O>(s, op, r, a, b, d_sc, sc, limbs), fn);
    if (sync) check(hipStreamSynchronize(s), fn);
}

template <class P>
void mont_batch(hipStream_t s, uint32_t gpu, P* pts, int n, bool sync, const char* fn) {
    if (n < 0) panic(fn, "invalid size n");
    if (n == 0) return;
    if (!pts) panic(fn, "null pointer argument");
    check(hipSetDevice((int)gpu), fn);
    check(gm::launch_point_mont_batch<typename OpsOf<P>::O>(s, pts, n, true), fn);
    if (sync) check(hipStreamSynchronize(s), fn);
}

// --------------------------------------------------------------- MSM

// Stream-ordered host finish of the asynchronous MSM. The window combine is a
// serial chain of about (scalar bits) doublings: one GPU lane runs each
// dependent 446-bit product at a few microseconds, a host core at a tenth of
// that, so the asynchronous entry points finish on the host without blocking
// the caller: the window sums are copied into a pinned slot, a host function
// enqueued on the caller's stream combines them and writes the ABI record into
// the slot, the record is copied to the caller's device pointer, and a last
// host function frees the slot. Every step is ordered on the stream; slots are
// reused, never freed before the process ends.
struct FinishSlot {
    uint8_t* buf = nullptr;  // pinned: window sums, then the record at res_off
    size_t cap = 0;
    std::atomic<bool> busy{false};
    int nw = 0, c = 0;
    bool proj = false;
    size_t res_off = 0;
};

std::mutex g_slot_mu;
std::vector<std::unique_ptr<FinishSlot>> g_slots;

FinishSlot* acquire_slot(size_t bytes, const char* fn) {
    std::lock_guard<std::mutex> lk(g_slot_mu);
    for (auto& sl : g_slots) {
        bool idle = false;
        if (sl->cap >= bytes && sl->busy.compare_exchange_strong(idle, true)) return sl.get();
    }
    auto sl = std::make_unique<FinishSlot>();
    check(hipHostMalloc(reinterpret_cast<void**>(&sl->buf), bytes, hipHostMallocPortable), fn);
    sl->cap = bytes;
    sl->busy = true;
    g_slots.push_back(std::move(sl));
    return g_slots.back().get();
}

template <class O>
void finish_slot(void* arg) {
    using HF = typename gmc::HostOf<O>::HF;
    FinishSlot* sl = static_cast<FinishSlot*>(arg);
    HF x, y;
    const bool finite = gmc::host_combine<HF>(x, y, reinterpret_cast<const uint32_t*>(sl->buf), sl->nw, sl->c);
    if (sl->proj)
        gmc::write_proj_abi<HF>(sl->buf + sl->res_off, finite, x, y);
    else
        gmc::write_affine_abi<HF>(sl->buf + sl->res_off, finite, x, y);
}

void release_slot(void* arg) { static_cast<FinishSlot*>(arg)->busy.store(false); }

template <class O>
bool plan_one(Plan& p, gm::MsmDesc& d, int n, int limbs) {
    gm::MsmSpec spec = {n, 0, 0, 0, 0};
    return gm::make_plan(p, &d, gmc::HostOf<O>::CURVE, limbs, 64 * limbs, &spec, 1, true);
}

template <class O>
size_t scratch_bytes(int n, int limbs) {
    if (n <= 0 || limbs < 1 || limbs > gm::MAX_SCALAR_LIMBS) return 0;
    Plan p;
    gm::MsmDesc d;
    if (!plan_one<O>(p, d, n, limbs)) return 0;
    return gm::al_size((size_t)n * sizeof(Aff<O>)) + p.total;
}

// Bytes from p to the end of the device allocation that holds it; 0 when the
// runtime cannot say. A caller that sizes its scratch the reference's way,
// (blocks + 1) * 16 projective points (reference tests/test_msm.cu:146-155),
// hands over far less than this pipeline's workspace (24 KB against 1.15 MB
// for G1 at n = 1000), so the MSM checks the span first and takes
// stream-ordered memory of its own when the buffer is too short or unknown,
// instead of writing past it.
size_t scratch_span(const void* p) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        return 0;
    }
    const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
    return (q >= b && q - b < size) ? size - (q - b) : 0;
}

// One MSM over device-resident ABI records. proj_out selects the ABI
// projective result (u64 entry points) or the affine record (multi-limb).
template <class O>
void msm_device(hipStream_t s, uint32_t gpu, void* d_result, const void* d_points, const uint64_t* d_scalars,
                int limbs, void* d_scratch, int n, bool proj_out, bool sync, const char* fn) {
    using HF = typename gmc::HostOf<O>::HF;
    using Lay = gm::AbiLayout<O>;
    if (n < 0) panic(fn, "invalid size n");
    if (!d_result) panic(fn, "null pointer argument");
    if (n == 0) {  // reference src/curve.cu:2539-2545 and 2766-2769: the result is O
        check(hipSetDevice((int)gpu), fn);
        if (proj_out)
            check(gm::launch_fill_u32(s, reinterpret_cast<uint32_t*>(d_result), Lay::PROJ_BYTES / 4, 0), fn);
        else
            check(gm::launch_point_op<O>(s, gm::PO_INF, d_result, nullptr, nullptr, nullptr, 0, 0), fn);
        if (sync) check(hipStreamSynchronize(s), fn);
        return;
    }
    if (!d_points || !d_scalars) panic(fn, "null pointer argument");
    if (limbs < 1 || limbs > gm::MAX_SCALAR_LIMBS) panic(fn, "scalar_limbs must be in [1, 8]");
    check(hipSetDevice((int)gpu), fn);
    Plan p;
    gm::MsmDesc d;
    if (!plan_one<O>(p, d, n, limbs)) panic(fn, "no pipeline plan for this size");
    const size_t b_aff = gm::al_size((size_t)n * sizeof(Aff<O>));
    uint8_t* scratch = reinterpret_cast<uint8_t*>(d_scratch);
    if (scratch && scratch_span(scratch) < b_aff + p.total) scratch = nullptr;  // too small: see scratch_span
    const bool own = scratch == nullptr;
    if (own) check(hipMallocAsync(reinterpret_cast<void**>(&scratch), b_aff + p.total, s), fn);
    Aff<O>* d_aff = reinterpret_cast<Aff<O>*>(scratch);
    uint8_t* ws = scratch + b_aff;
    check(gm::launch_put_desc(s, reinterpret_cast<gm::MsmDesc*>(ws + p.o_desc), d), fn);
    check(gm::Engine<O>::import_aff(s, d_aff, d_points, n, false), fn);
    check(gm::Engine<O>::run(s, p, d_aff, d_scalars, ws, nullptr), fn);
    if (!sync) {
        // stream-ordered host finish (FinishSlot): nothing here waits
        const size_t wb = (size_t)d.nw * sizeof(Xyzz<O>), res_off = gm::al_size(wb);
        FinishSlot* sl = acquire_slot(res_off + Lay::PROJ_BYTES, fn);
        sl->nw = (int)d.nw;
        sl->c = (int)d.c;
        sl->proj = proj_out;
        sl->res_off = res_off;
        check(hipMemcpyAsync(sl->buf, ws + p.o_wsum, wb, hipMemcpyDeviceToHost, s), fn);
        if (own) check(hipFreeAsync(scratch, s), fn);
        check(hipLaunchHostFunc(s, finish_slot<O>, sl), fn);
        check(hipMemcpyAsync(d_result, sl->buf + res_off, proj_out ? Lay::PROJ_BYTES : Lay::AFF_BYTES,
                             hipMemcpyHostToDevice, s),
              fn);
        check(hipLaunchHostFunc(s, release_slot, sl), fn);
        return;
    }
    // synchronous: window sums to the host, Horner there, the record back
    std::vector<uint8_t> h((size_t)d.nw * sizeof(Xyzz<O>));
    check(hipMemcpyAsync(h.data(), ws + p.o_wsum, h.size(), hipMemcpyDeviceToHost, s), fn);
    check(hipStreamSynchronize(s), fn);
    HF x, y;
    const bool finite = gmc::host_combine<HF>(x, y, reinterpret_cast<const uint32_t*>(h.data()), (int)d.nw, (int)d.c);
    uint8_t rec[Lay::PROJ_BYTES];
    size_t bytes;
    if (proj_out) {
        gmc::write_proj_abi<HF>(rec, finite, x, y);
        bytes = Lay::PROJ_BYTES;
    } else {
        gmc::write_affine_abi<HF>(rec, finite, x, y);
        bytes = Lay::AFF_BYTES;
    }
    check(hipMemcpyAsync(d_result, rec, bytes, hipMemcpyHostToDevice, s), fn);
    if (own) check(hipFreeAsync(scratch, s), fn);
    check(hipStreamSynchronize(s), fn);
}

// ------------------------------------------------------ field elements

template <class T>
void field_dev(hipStream_t s, uint32_t gpu, int op, T* c, const T* a, const T* b, int n, uint64_t arg, int* iout,
               bool sync, const char* fn) {
    check(hipSetDevice((int)gpu), fn);
    check(gm::launch_field_op<typename OpsOf<T>::O>(s, op, c, a, b, n, arg, iout), fn);
    if (sync) check(hipStreamSynchronize(s), fn);
}

// Host-pointer form: stream-ordered device copies of the operands (c is
// copied in as well, for cmov), one launch, results back, synchronised.
// Returns the integer output (comparison or flag) when the op has one.
template <class T>
int field_host(hipStream_t s, uint32_t gpu, int op, T* c, const T* a, const T* b, int n, uint64_t arg, bool c_in,
               bool has_iout, const char* fn) {
    if (n < 0) panic(fn, "invalid size n");
    if (n == 0) return 0;
    check(hipSetDevice((int)gpu), fn);
    const size_t eb = (size_t)n * sizeof(T);
    uint8_t* d = nullptr;
    const size_t total = 3 * gm::al_size(eb) + gm::al_size((size_t)n * sizeof(int));
    check(hipMallocAsync(reinterpret_cast<void**>(&d), total, s), fn);
    T* dc = reinterpret_cast<T*>(d);
    T* da = reinterpret_cast<T*>(d + gm::al_size(eb));
    T* db = reinterpret_cast<T*>(d + 2 * gm::al_size(eb));
    int* di = reinterpret_cast<int*>(d + 3 * gm::al_size(eb));
    if (a) check(hipMemcpyAsync(da, a, eb, hipMemcpyHostToDevice, s), fn);
    if (b) check(hipMemcpyAsync(db, b, eb, hipMemcpyHostToDevice, s), fn);
    if (c_in) check(hipMemcpyAsync(dc, c, eb, hipMemcpyHostToDevice, s), fn);
    check(gm::launch_field_op<typename OpsOf<T>::O>(s, op, dc, a ? da : nullptr, b ? db : nullptr, n, arg, di), fn);
    std::vector<int> iv(has_iout ? (size_t)n : 0);
    if (c) check(hipMemcpyAsync(c, dc, eb, hipMemcpyDeviceToHost, s), fn);
    if (has_iout) check(hipMemcpyAsync(iv.data(), di, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s), fn);
    check(hipFreeAsync(d, s), fn);
    check(hipStreamSynchronize(s), fn);
    return has_iout ? iv[0] : 0;
}

void need(bool ok, const char* fn) {
    if (!ok) panic(fn, "null pointer argument");
}

// The stream-less Fp2 functions run on the current device's null stream.
uint32_t current_gpu(const char* fn) {
    int dev = 0;
    check(hipGetDevice(&dev), fn);
    return (uint32_t)dev;
}

}  // namespace

// ===================================================== include/curve.h

template <typename P>
void point_add_async(hipStream_t s, uint32_t g, P* r, const P* a, const P* b) {
    need(r && a && b, "point_add_async");
    point_op<P>(s, g, gm::PO_ADD, r, a, b, nullptr, 0, 0, false, "point_add_async");
}
template <typename P>
void point_add(hipStream_t s, uint32_t g, P* r, const P* a, const P* b) {
    need(r && a && b, "point_add");
    point_op<P>(s, g, gm::PO_ADD, r, a, b, nullptr, 0, 0, true, "point_add");
}
template <typename P>
void point_double_async(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_double_async");
    point_op<P>(s, g, gm::PO_DOUBLE, r, a, nullptr, nullptr, 0, 0, false, "point_double_async");
}
template <typename P>
void point_double(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_double");
    point_op<P>(s, g, gm::PO_DOUBLE, r, a, nullptr, nullptr, 0, 0, true, "point_double");
}
template <typename P>
void point_neg_async(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_neg_async");
    point_op<P>(s, g, gm::PO_NEG, r, a, nullptr, nullptr, 0, 0, false, "point_neg_async");
}
template <typename P>
void point_neg(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_neg");
    point_op<P>(s, g, gm::PO_NEG, r, a, nullptr, nullptr, 0, 0, true, "point_neg");
}
template <typename P>
void point_scalar_mul_u64_async(hipStream_t s, uint32_t g, P* r, const P* a, uint64_t k) {
    need(r && a, "point_scalar_mul_u64_async");
    point_op<P>(s, g, gm::PO_SMUL, r, a, nullptr, nullptr, k, 1, false, "point_scalar_mul_u64_async");
}
template <typename P>
void point_scalar_mul_u64(hipStream_t s, uint32_t g, P* r, const P* a, uint64_t k) {
    need(r && a, "point_scalar_mul_u64");
    point_op<P>(s, g, gm::PO_SMUL, r, a, nullptr, nullptr, k, 1, true, "point_scalar_mul_u64");
}
template <typename P>
void point_scalar_mul_async(hipStream_t s, uint32_t g, P* r, const P* a, const uint64_t* k, int limbs) {
    need(r && a && k, "point_scalar_mul_async");
    point_op<P>(s, g, gm::PO_SMUL, r, a, nullptr, k, 0, limbs, false, "point_scalar_mul_async");
}
template <typename P>
void point_scalar_mul(hipStream_t s, uint32_t g, P* r, const P* a, const uint64_t* k, int limbs) {
    need(r && a && k, "point_scalar_mul");
    point_op<P>(s, g, gm::PO_SMUL, r, a, nullptr, k, 0, limbs, true, "point_scalar_mul");
}
template <typename P>
void point_at_infinity_async(hipStream_t s, uint32_t g, P* r) {
    need(r, "point_at_infinity_async");
    point_op<P>(s, g, gm::PO_INF, r, nullptr, nullptr, nullptr, 0, 0, false, "point_at_infinity_async");
}
template <typename P>
void point_at_infinity(hipStream_t s, uint32_t g, P* r) {
    need(r, "point_at_infinity");
    point_op<P>(s, g, gm::PO_INF, r, nullptr, nullptr, nullptr, 0, 0, true, "point_at_infinity");
}
template <typename P>
void point_to_montgomery_async(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_to_montgomery_async");
    point_op<P>(s, g, gm::PO_TO_MONT, r, a, nullptr, nullptr, 0, 0, false, "point_to_montgomery_async");
}
template <typename P>
void point_to_montgomery(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_to_montgomery");
    point_op<P>(s, g, gm::PO_TO_MONT, r, a, nullptr, nullptr, 0, 0, true, "point_to_montgomery");
}
template <typename P>
void point_from_montgomery_async(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_from_montgomery_async");
    point_op<P>(s, g, gm::PO_FROM_MONT, r, a, nullptr, nullptr, 0, 0, false, "point_from_montgomery_async");
}
template <typename P>
void point_from_montgomery(hipStream_t s, uint32_t g, P* r, const P* a) {
    need(r && a, "point_from_montgomery");
    point_op<P>(s, g, gm::PO_FROM_MONT, r, a, nullptr, nullptr, 0, 0, true, "point_from_montgomery");
}
template <typename P>
void point_to_montgomery_batch_async(hipStream_t s, uint32_t g, P* pts, int n) {
    mont_batch<P>(s, g, pts, n, false, "point_to_montgomery_batch_async");
}
template <typename P>
void point_to_montgomery_batch(hipStream_t s, uint32_t g, P* pts, int n) {
    mont_batch<P>(s, g, pts, n, true, "point_to_montgomery_batch");
}
template <typename P>
size_t point_msm_scratch_bytes(int n, int limbs) {
    return scratch_bytes<typename OpsOf<P>::O>(n, limbs);
}
template <typename P>
void point_msm_async(hipStream_t s, uint32_t g, P* r, const P* pts, const uint64_t* sc, int limbs, P* scratch, int n) {
    msm_device<typename OpsOf<P>::O>(s, g, r, pts, sc, limbs, scratch, n, false, false, "point_msm_async");
}
template <typename P>
void point_msm(hipStream_t s, uint32_t g, P* r, const P* pts, const uint64_t* sc, int limbs, P* scratch, int n) {
    msm_device<typename OpsOf<P>::O>(s, g, r, pts, sc, limbs, scratch, n, false, true, "point_msm");
}

// The PointType instantiations of reference src/curve.cu:2832-2876.
#define GM_POINT_INST(T)                                                                                    \
    template void point_add_async<T>(hipStream_t, uint32_t, T*, const T*, const T*);                       \
    template void point_add<T>(hipStream_t, uint32_t, T*, const T*, const T*);                             \
    template void point_double_async<T>(hipStream_t, uint32_t, T*, const T*);                              \
    template void point_double<T>(hipStream_t, uint32_t, T*, const T*);                                    \
    template void point_neg_async<T>(hipStream_t, uint32_t, T*, const T*);                                 \
    template void point_neg<T>(hipStream_t, uint32_t, T*, const T*);                                       \
    template void point_at_infinity_async<T>(hipStream_t, uint32_t, T*);                                   \
    template void point_at_infinity<T>(hipStream_t, uint32_t, T*);                                         \
    template void point_to_montgomery_async<T>(hipStream_t, uint32_t, T*, const T*);                       \
    template void point_to_montgomery<T>(hipStream_t, uint32_t, T*, const T*);                             \
    template void point_from_montgomery_async<T>(hipStream_t, uint32_t, T*, const T*);                     \
    template void point_from_montgomery<T>(hipStream_t, uint32_t, T*, const T*);                           \
    template void point_scalar_mul_u64_async<T>(hipStream_t, uint32_t, T*, const T*, uint64_t);            \
    template void point_scalar_mul_u64<T>(hipStream_t, uint32_t, T*, const T*, uint64_t);                  \
    template void point_scalar_mul_async<T>(hipStream_t, uint32_t, T*, const T*, const uint64_t*, int);    \
    template void point_scalar_mul<T>(hipStream_t, uint32_t, T*, const T*, const uint64_t*, int);          \
    template void point_to_montgomery_batch_async<T>(hipStream_t, uint32_t, T*, int);                      \
    template void point_to_montgomery_batch<T>(hipStream_t, uint32_t, T*, int);                            \
    template size_t point_msm_scratch_bytes<T>(int, int);                                                  \
    template void point_msm_async<T>(hipStream_t, uint32_t, T*, const T*, const uint64_t*, int, T*, int);  \
    template void point_msm<T>(hipStream_t, uint32_t, T*, const T*, const uint64_t*, int, T*, int);

GM_POINT_INST(G1Point)
GM_POINT_INST(G2Point)
#undef GM_POINT_INST

size_t point_msm_u64_scratch_bytes_g1(int n) { return scratch_bytes<FpOps>(n, 1); }
size_t point_msm_u64_scratch_bytes_g2(int n) { return scratch_bytes<Fp2Ops>(n, 1); }

void point_msm_u64_async_g1(hipStream_t s, uint32_t g, G1ProjectivePoint* r, const G1Point* pts, const uint64_t* sc,
                            G1ProjectivePoint* scratch, int n) {
    msm_device<FpOps>(s, g, r, pts, sc, 1, scratch, n, true, false, "point_msm_u64_async_g1");
}
void point_msm_u64_g1(hipStream_t s, uint32_t g, G1ProjectivePoint* r, const G1Point* pts, const uint64_t* sc,
                      G1ProjectivePoint* scratch, int n) {
    msm_device<FpOps>(s, g, r, pts, sc, 1, scratch, n, true, true, "point_msm_u64_g1");
}
void point_msm_u64_async_g2(hipStream_t s, uint32_t g, G2ProjectivePoint* r, const G2Point* pts, const uint64_t* sc,
                            G2ProjectivePoint* scratch, int n) {
    msm_device<Fp2Ops>(s, g, r, pts, sc, 1, scratch, n, true, false, "point_msm_u64_async_g2");
}
void point_msm_u64_g2(hipStream_t s, uint32_t g, G2ProjectivePoint* r, const G2Point* pts, const uint64_t* sc,
                      G2ProjectivePoint* scratch, int n) {
    msm_device<Fp2Ops>(s, g, r, pts, sc, 1, scratch, n, true, true, "point_msm_u64_g2");
}

// ================================================= include/fp.h, fp2.h

#define GM_FIELD_BIN(T, name, op)                                                   \
    void name##_async(hipStream_t s, uint32_t g, T* c, const T* a, const T* b) {    \
        need(c && a && b, #name "_async");                                          \
        field_dev<T>(s, g, op, c, a, b, 1, 0, nullptr, false, #name "_async");      \
    }                                                                               \
    void name(hipStream_t s, uint32_t g, T* c, const T* a, const T* b) {            \
        need(c && a && b, #name);                                                   \
        field_dev<T>(s, g, op, c, a, b, 1, 0, nullptr, true, #name);                \
    }
#define GM_FIELD_UN(T, name, op)                                                    \
    void name##_async(hipStream_t s, uint32_t g, T* c, const T* a) {                \
        need(c && a, #name "_async");                                               \
        field_dev<T>(s, g, op, c, a, nullptr, 1, 0, nullptr, false, #name "_async"); \
    }                                                                               \
    void name(hipStream_t s, uint32_t g, T* c, const T* a) {                        \
        need(c && a, #name);                                                        \
        field_dev<T>(s, g, op, c, a, nullptr, 1, 0, nullptr, true, #name);          \
    }
#define GM_FIELD_SET(T, name, op)                                                   \
    void name##_async(hipStream_t s, uint32_t g, T* c) {                            \
        need(c, #name "_async");                                                    \
        field_dev<T>(s, g, op, c, (const T*)nullptr, (const T*)nullptr, 1, 0, nullptr, false, #name "_async"); \
    }                                                                               \
    void name(hipStream_t s, uint32_t g, T* c) {                                    \
        need(c, #name);                                                             \
        field_dev<T>(s, g, op, c, (const T*)nullptr, (const T*)nullptr, 1, 0, nullptr, true, #name); \
    }

GM_FIELD_BIN(Fp, fp_add, gm::FO_ADD)
GM_FIELD_BIN(Fp, fp_sub, gm::FO_SUB)
GM_FIELD_BIN(Fp, fp_mul, gm::FO_MUL)
GM_FIELD_BIN(Fp, fp_div, gm::FO_DIV)
GM_FIELD_BIN(Fp, fp_mont_mul, gm::FO_MONT_MUL)
GM_FIELD_UN(Fp, fp_neg, gm::FO_NEG)
GM_FIELD_UN(Fp, fp_inv, gm::FO_INV)
GM_FIELD_UN(Fp, fp_copy, gm::FO_COPY)
GM_FIELD_UN(Fp, fp_to_montgomery, gm::FO_TO_MONT)
GM_FIELD_UN(Fp, fp_from_montgomery, gm::FO_FROM_MONT)
GM_FIELD_SET(Fp, fp_zero, gm::FO_ZERO)
GM_FIELD_SET(Fp, fp_one, gm::FO_ONE)

GM_FIELD_BIN(Fp2, fp2_add, gm::FO_ADD)
GM_FIELD_BIN(Fp2, fp2_sub, gm::FO_SUB)
GM_FIELD_BIN(Fp2, fp2_mul, gm::FO_MUL)
GM_FIELD_BIN(Fp2, fp2_div, gm::FO_DIV)
GM_FIELD_UN(Fp2, fp2_square, gm::FO_SQUARE)
GM_FIELD_UN(Fp2, fp2_neg, gm::FO_NEG)
GM_FIELD_UN(Fp2, fp2_conjugate, gm::FO_CONJ)
GM_FIELD_UN(Fp2, fp2_inv, gm::FO_INV)
GM_FIELD_UN(Fp2, fp2_copy, gm::FO_COPY)

#undef GM_FIELD_BIN
#undef GM_FIELD_UN
#undef GM_FIELD_SET

// ======================================== include/fp_kernels.h, fp2_kernels.h

void fp_add_array_host(hipStream_t s, uint32_t g, Fp* c, const Fp* a, const Fp* b, int n) {
    if (n > 0) need(c && a && b, "fp_add_array_host");
    field_host<Fp>(s, g, gm::FO_ADD, c, a, b, n, 0, false, false, "fp_add_array_host");
}
void fp_mul_array_host(hipStream_t s, uint32_t g, Fp* c, const Fp* a, const Fp* b, int n) {
    if (n > 0) need(c && a && b, "fp_mul_array_host");
    field_host<Fp>(s, g, gm::FO_MUL, c, a, b, n, 0, false, false, "fp_mul_array_host");
}
void fp_add_array_device(hipStream_t s, uint32_t g, Fp* c, const Fp* a, const Fp* b, int n) {
    if (n < 0) panic("fp_add_array_device", "invalid size n");
    if (n == 0) return;
    need(c && a && b, "fp_add_array_device");
    field_dev<Fp>(s, g, gm::FO_ADD, c, a, b, n, 0, nullptr, false, "fp_add_array_device");
}
void fp_mul_array_device(hipStream_t s, uint32_t g, Fp* c, const Fp* a, const Fp* b, int n) {
    if (n < 0) panic("fp_mul_array_device", "invalid size n");
    if (n == 0) return;
    need(c && a && b, "fp_mul_array_device");
    field_dev<Fp>(s, g, gm::FO_MUL, c, a, b, n, 0, nullptr, false, "fp_mul_array_device");
}

#define GM_FP_GPU_BIN(name, op)                                                   \
    void name(hipStream_t s, uint32_t g, Fp* r, const Fp* a, const Fp* b) {       \
        need(r && a && b, #name);                                                 \
        field_host<Fp>(s, g, op, r, a, b, 1, 0, false, false, #name);             \
    }
#define GM_FP_GPU_UN(name, op)                                                    \
    void name(hipStream_t s, uint32_t g, Fp* r, const Fp* a) {                    \
        need(r && a, #name);                                                      \
        field_host<Fp>(s, g, op, r, a, (const Fp*)nullptr, 1, 0, false, false, #name); \
    }
GM_FP_GPU_BIN(fp_add_gpu, gm::FO_ADD)
GM_FP_GPU_BIN(fp_sub_gpu, gm::FO_SUB)
GM_FP_GPU_BIN(fp_mul_gpu, gm::FO_MUL)
GM_FP_GPU_BIN(fp_div_gpu, gm::FO_DIV)
GM_FP_GPU_BIN(fp_mont_mul_gpu, gm::FO_MONT_MUL)
GM_FP_GPU_UN(fp_neg_gpu, gm::FO_NEG)
GM_FP_GPU_UN(fp_inv_gpu, gm::FO_INV)
GM_FP_GPU_UN(fp_to_montgomery_gpu, gm::FO_TO_MONT)
GM_FP_GPU_UN(fp_from_montgomery_gpu, gm::FO_FROM_MONT)
GM_FP_GPU_UN(fp_copy_gpu, gm::FO_COPY)
#undef GM_FP_GPU_BIN
#undef GM_FP_GPU_UN

int fp_cmp_gpu(hipStream_t s, uint32_t g, const Fp* a, const Fp* b) {
    need(a && b, "fp_cmp_gpu");
    return field_host<Fp>(s, g, gm::FO_CMP, (Fp*)nullptr, a, b, 1, 0, false, true, "fp_cmp_gpu");
}
bool fp_is_zero_gpu(hipStream_t s, uint32_t g, const Fp* a) {
    need(a, "fp_is_zero_gpu");
    return field_host<Fp>(s, g, gm::FO_IS_ZERO, (Fp*)nullptr, a, (const Fp*)nullptr, 1, 0, false, true,
                          "fp_is_zero_gpu") != 0;
}
bool fp_is_one_gpu(hipStream_t s, uint32_t g, const Fp* a) {
    need(a, "fp_is_one_gpu");
    return field_host<Fp>(s, g, gm::FO_IS_ONE, (Fp*)nullptr, a, (const Fp*)nullptr, 1, 0, false, true,
                          "fp_is_one_gpu") != 0;
}
void fp_cmov_gpu(hipStream_t s, uint32_t g, Fp* r, const Fp* src, uint64_t cond) {
    need(r && src, "fp_cmov_gpu");
    field_host<Fp>(s, g, gm::FO_CMOV, r, src, (const Fp*)nullptr, 1, cond, true, false, "fp_cmov_gpu");
}
bool fp_sqrt_gpu(hipStream_t s, uint32_t g, Fp* r, const Fp* a) {
    need(r && a, "fp_sqrt_gpu");
    return field_host<Fp>(s, g, gm::FO_SQRT, r, a, (const Fp*)nullptr, 1, 0, false, true, "fp_sqrt_gpu") != 0;
}
bool fp_is_quadratic_residue_gpu(hipStream_t s, uint32_t g, const Fp* a) {
    need(a, "fp_is_quadratic_residue_gpu");
    return field_host<Fp>(s, g, gm::FO_IS_QR, (Fp*)nullptr, a, (const Fp*)nullptr, 1, 0, false, true,
                          "fp_is_quadratic_residue_gpu") != 0;
}
void fp_pow_u64_gpu(hipStream_t s, uint32_t g, Fp* r, const Fp* base, uint64_t e) {
    need(r && base, "fp_pow_u64_gpu");
    field_host<Fp>(s, g, gm::FO_POW_U64, r, base, (const Fp*)nullptr, 1, e, false, false, "fp_pow_u64_gpu");
}

void fp2_add_array_host(hipStream_t s, uint32_t g, Fp2* c, const Fp2* a, const Fp2* b, int n) {
    if (n > 0) need(c && a && b, "fp2_add_array_host");
    field_host<Fp2>(s, g, gm::FO_ADD, c, a, b, n, 0, false, false, "fp2_add_array_host");
}
void fp2_mul_array_host(hipStream_t s, uint32_t g, Fp2* c, const Fp2* a, const Fp2* b, int n) {
    if (n > 0) need(c && a && b, "fp2_mul_array_host");
    field_host<Fp2>(s, g, gm::FO_MUL, c, a, b, n, 0, false, false, "fp2_mul_array_host");
}

#define GM_FP2_GPU_BIN(name, op)                                                                  \
    void name(Fp2* r, const Fp2* a, const Fp2* b) {                                               \
        need(r && a && b, #name);                                                                 \
        field_host<Fp2>(nullptr, current_gpu(#name), op, r, a, b, 1, 0, false, false, #name);     \
    }
#define GM_FP2_GPU_UN(name, op)                                                                   \
    void name(Fp2* r, const Fp2* a) {                                                             \
        need(r && a, #name);                                                                      \
        field_host<Fp2>(nullptr, current_gpu(#name), op, r, a, (const Fp2*)nullptr, 1, 0, false, false, #name); \
    }
GM_FP2_GPU_BIN(fp2_add_gpu, gm::FO_ADD)
GM_FP2_GPU_BIN(fp2_sub_gpu, gm::FO_SUB)
GM_FP2_GPU_BIN(fp2_mul_gpu, gm::FO_MUL)
GM_FP2_GPU_BIN(fp2_div_gpu, gm::FO_DIV)
GM_FP2_GPU_UN(fp2_neg_gpu, gm::FO_NEG)
GM_FP2_GPU_UN(fp2_conjugate_gpu, gm::FO_CONJ)
GM_FP2_GPU_UN(fp2_square_gpu, gm::FO_SQUARE)
GM_FP2_GPU_UN(fp2_inv_gpu, gm::FO_INV)
GM_FP2_GPU_UN(fp2_mul_by_i_gpu, gm::FO_MUL_BY_I)
GM_FP2_GPU_UN(fp2_frobenius_gpu, gm::FO_CONJ)
GM_FP2_GPU_UN(fp2_copy_gpu, gm::FO_COPY)
#undef GM_FP2_GPU_BIN
#undef GM_FP2_GPU_UN

int fp2_cmp_gpu(const Fp2* a, const Fp2* b) {
    need(a && b, "fp2_cmp_gpu");
    return field_host<Fp2>(nullptr, current_gpu("fp2_cmp_gpu"), gm::FO_CMP, (Fp2*)nullptr, a, b, 1, 0, false, true,
                           "fp2_cmp_gpu");
}
bool fp2_is_zero_gpu(const Fp2* a) {
    need(a, "fp2_is_zero_gpu");
    return field_host<Fp2>(nullptr, current_gpu("fp2_is_zero_gpu"), gm::FO_IS_ZERO, (Fp2*)nullptr, a,
                           (const Fp2*)nullptr, 1, 0, false, true, "fp2_is_zero_gpu") != 0;
}
bool fp2_is_one_gpu(const Fp2* a) {
    need(a, "fp2_is_one_gpu");
    return field_host<Fp2>(nullptr, current_gpu("fp2_is_one_gpu"), gm::FO_IS_ONE, (Fp2*)nullptr, a,
                           (const Fp2*)nullptr, 1, 0, false, true, "fp2_is_one_gpu") != 0;
}
void fp2_cmov_gpu(Fp2* r, const Fp2* src, uint64_t cond) {
    need(r && src, "fp2_cmov_gpu");
    field_host<Fp2>(nullptr, current_gpu("fp2_cmov_gpu"), gm::FO_CMOV, r, src, (const Fp2*)nullptr, 1, cond, true,
                    false, "fp2_cmov_gpu");
}
