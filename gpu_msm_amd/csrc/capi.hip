// C ABI of the MSM library: the reference's rust_wrapper.h surface
// (reference rust-api/src/c_wrapper.cu) plus the extensions of
// include/gpu_msm_ext.h. Host code: per-device context (stream, cached
// workspace, pinned staging), plan selection, the host window combine.
#include <hipcub/hipcub.hpp>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/gpu_msm_ext.h"
#include "../../include/rust_wrapper.h"
#include "engine.hpp"
#include "format.hpp"
#include "host_combine.hpp"
#include "host_field.hpp"
#include "resident.hpp"
#include "staging.hpp"

#define GM_TRY_HOST(x)                   \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

using gm::Aff;
using gm::Fp2Ops;
using gm::FpOps;
using gm::Plan;
using gm::Xyzz;
using gmc::AffResult;
using gmc::dev_elem_to_host;
using gmc::get_elem;
using gmc::HostOf;
using gmc::put_elem;
using gmc::words_of;
using gmc::write_affine_abi;
using gmc::write_proj_abi;
using gmc::host_combine;

// ------------------------------------------------------------------ context

namespace {

struct DevCtx {
    int dev = 0;
    hipStream_t stream = nullptr;
    uint8_t* ws = nullptr;
    size_t cap = 0;
    uint8_t* pinned = nullptr;  // host staging for window sums / results
    size_t pinned_cap = 0;
    gmc::Staging stage;
    std::mutex mu;
    gm::Prof prof{};
    bool prof_valid = false;
    // last MSM: window, windows, (window, bucket) entries accumulated
    int last_c = 0, last_nw = 0;
    long long last_entries = 0;
};

std::mutex g_ctx_mu;
DevCtx* g_ctx[64];

DevCtx* get_ctx(uint32_t gpu) {
    if (gpu >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (g_ctx[gpu]) return g_ctx[gpu];
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || (int)gpu >= count) return nullptr;
    if (hipSetDevice((int)gpu) != hipSuccess) return nullptr;
    DevCtx* c = new DevCtx();
    c->dev = (int)gpu;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    for (int i = 0; i < gm::ST_COUNT; ++i)
        if (hipEventCreate(&c->prof.ev[i]) != hipSuccess) {
            delete c;
            return nullptr;
        }
    c->prof.on = false;
    g_ctx[gpu] = c;
    return c;
}

hipError_t ensure_ws(DevCtx* c, size_t bytes) {
    if (bytes <= c->cap) return hipSuccess;
    if (c->ws) {
        hipStreamSynchronize(c->stream);
        hipFree(c->ws);
        c->ws = nullptr;
        c->cap = 0;
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&c->ws, want);
    if (e != hipSuccess) {
        c->ws = nullptr;
        return e;
    }
    c->cap = want;
    return hipSuccess;
}

hipError_t ensure_pinned(DevCtx* c, size_t bytes) {
    if (bytes <= c->pinned_cap) return hipSuccess;
    if (c->pinned) {
        hipStreamSynchronize(c->stream);
        hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_cap = 0;
    }
    hipError_t e = hipHostMalloc((void**)&c->pinned, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        c->pinned = nullptr;
        return e;
    }
    c->pinned_cap = bytes;
    return hipSuccess;
}

// Host -> device copies of pageable caller buffers on c->stream through the
// context's pinned staging. Caller holds c->mu.
hipError_t h2d_staged(DevCtx* c, const gmc::H2D* cp, int ncp) {
    return gmc::h2d_staged(c->stage, c->dev, c->stream, cp, ncp);
}

hipError_t h2d_staged(DevCtx* c, void* dst, const void* src, size_t bytes) {
    const gmc::H2D one = {dst, src, bytes};
    return h2d_staged(c, &one, 1);
}

template <class HF>
void host_combine_many(std::vector<AffResult<HF>>& out, const uint32_t* raw, const gm::MsmDesc* desc, int count);

// Runs the pipeline for a batch of MSMs over resident bases/scalars and
// finishes each with a host (or device) window combine. Caller holds c->mu.
template <class O>
int run_batch_locked(DevCtx* c, const Aff<O>* d_bases, const uint64_t* d_scalars, int L, int bits,
                     const gm::MsmSpec* specs, int count, bool device_combine,
                     std::vector<AffResult<typename HostOf<O>::HF>>& out, uint8_t* ws_override = nullptr) {
    using HF = typename HostOf<O>::HF;
    std::vector<gm::MsmDesc> desc((size_t)count);
    Plan p;
    if (!gm::make_plan(p, desc.data(), HostOf<O>::CURVE, L, bits, specs, count)) return -1;
    uint8_t* ws = ws_override;
    if (!ws) {
        if (ensure_ws(c, p.total) != hipSuccess) return -3;
        ws = c->ws;
    }
    const size_t desc_b = (size_t)count * sizeof(gm::MsmDesc);
    const size_t res_b = device_combine ? (size_t)count * sizeof(Aff<O>) : (size_t)p.win_total * sizeof(Xyzz<O>);
    // pinned: the descriptors, then the error words and the results
    // ERR_BYTES apart, as the workspace holds the error words and the window
    // sums, so that one copy brings both back
    const size_t o_flags = gm::al_size(desc_b), o_res = o_flags + gm::ERR_BYTES;
    if (ensure_pinned(c, o_res + gm::al_size(res_b)) != hipSuccess) return -3;
    uint8_t* h = c->pinned;
    memcpy(h, desc.data(), desc_b);
    if (hipMemcpyAsync(ws + p.o_desc, h, desc_b, hipMemcpyHostToDevice, c->stream) != hipSuccess) return -4;
    if (gm::Engine<O>::run(c->stream, p, d_bases, d_scalars, ws, &c->prof) != hipSuccess) return -4;
    c->prof_valid = c->prof.on;
    if (device_combine) {
        if (gm::Engine<O>::device_combine(c->stream, p, ws) != hipSuccess) return -4;
        if (hipMemcpyAsync(h + o_flags, ws + p.o_err, 16, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return -4;
        if (hipMemcpyAsync(h + o_res, ws + p.o_out, res_b, hipMemcpyDeviceToHost, c->stream) != hipSuccess) return -4;
    } else if (hipMemcpyAsync(h + o_flags, ws + p.o_err, gm::ERR_BYTES + res_b, hipMemcpyDeviceToHost, c->stream) !=
               hipSuccess) {
        return -4;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -4;
    uint32_t err, entries;
    memcpy(&err, h + o_flags, 4);
    memcpy(&entries, h + o_flags + 8, 4);
    if (err) return -5;
    c->last_c = (int)desc[0].c;
    c->last_nw = (int)desc[0].nw;
    c->last_entries = entries;
    out.resize((size_t)count);
    constexpr int W = words_of<HF>();
    if (device_combine) {
        for (int j = 0; j < count; ++j) {
            const uint32_t* a = reinterpret_cast<const uint32_t*>(h + o_res + (size_t)j * sizeof(Aff<O>));
            out[j].finite = a[15] != gm::INF_MARK;
            if (out[j].finite) {
                dev_elem_to_host(out[j].x, a);
                dev_elem_to_host(out[j].y, a + W);
            }
        }
    } else {
        host_combine_many<HF>(out, reinterpret_cast<const uint32_t*>(h + o_res), desc.data(), count);
    }
    return 0;
}

// Host-pointer MSM: upload, convert, run, finish. Used by the reference-ABI
// wrappers.
template <class O>
int msm_host_ptrs(void* result_proj, const void* points, const uint64_t* scalars, int L, int bits, int n,
                  uint32_t gpu, bool to_mont) {
    using HF = typename HostOf<O>::HF;
    using Lay = gm::AbiLayout<O>;
    if (n <= 0 || !result_proj || !points || !scalars || L < 1 || L > 8) return -1;
    DevCtx* c = get_ctx(gpu);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->dev) != hipSuccess) return -2;
    gm::MsmSpec spec = {n, 0, 0, 0, 0};
    gm::MsmDesc d1;
    Plan p;
    if (!gm::make_plan(p, &d1, HostOf<O>::CURVE, L, bits, &spec, 1)) return -1;
    const size_t b_abi = gm::al_size((size_t)n * Lay::AFF_BYTES);
    const size_t b_aff = gm::al_size((size_t)n * sizeof(Aff<O>));
    const size_t b_sc = gm::al_size((size_t)n * L * 8);
    if (ensure_ws(c, b_abi + b_aff + b_sc + p.total) != hipSuccess) return -3;
    uint8_t* d_abi = c->ws;
    Aff<O>* d_aff = reinterpret_cast<Aff<O>*>(c->ws + b_abi);
    uint64_t* d_sc = reinterpret_cast<uint64_t*>(c->ws + b_abi + b_aff);
    uint8_t* ws = c->ws + b_abi + b_aff + b_sc;
    const gmc::H2D cp[2] = {{d_abi, points, (size_t)n * Lay::AFF_BYTES}, {d_sc, scalars, (size_t)n * L * 8}};
    if (h2d_staged(c, cp, 2) != hipSuccess) return -4;
    if (gm::Engine<O>::import_aff(c->stream, d_aff, d_abi, n, to_mont) != hipSuccess) return -4;
    std::vector<AffResult<HF>> res;
    int rc = run_batch_locked<O>(c, d_aff, d_sc, L, bits, &spec, 1, false, res, ws);
    if (rc) return rc;
    write_proj_abi<HF>(result_proj, res[0].finite, res[0].x, res[0].y);
    return 0;
}

}  // namespace

// ------------------------------------------------------------ resident data


namespace {

template <class O>
Aff<O> generator_dev();
template <>
Aff<FpOps> generator_dev<FpOps>() {
    // reference src/curve.cu:786-796 (standard form)
    static const gmh::Fp gx = {{0x3bf9166c8236f4faULL, 0x8bc02b7cbe6a9e8dULL, 0x11c1e56b3e4bc80bULL,
                                0x6b20d782901a6f62ULL, 0x2ce8c34265bf3841ULL, 0x11b73d3d76ae9851ULL,
                                0x326ed6bd777fc6a3ULL}};
    static const gmh::Fp gy = {{0xfe6f792612016b30ULL, 0x22db0ce6034a9db9ULL, 0xb9093f32002756daULL,
                                0x39d7f424b6660204ULL, 0xf843c947aa57f571ULL, 0xd6d62d244e413636ULL,
                                0x1a7caf4a4d3887a6ULL}};
    Aff<FpOps> g;
    gmh::Fp t;
    gmh::to_mont(t, gx);
    gmh::to_limbs28(g.x.v, t);
    gmh::to_mont(t, gy);
    gmh::to_limbs28(g.y.v, t);
    return g;
}
template <>
Aff<Fp2Ops> generator_dev<Fp2Ops>() {
    // reference src/curve.cu:798-814 (standard form)
    static const gmh::Fp c[4] = {
        {{0x0e529ee4dce9991dULL, 0xd6ebaf149094f1ccULL, 0x043c6bf16312d638ULL, 0x062b61439640e885ULL,
          0x18dad8ed784dd225ULL, 0xa57c0038441f7d15ULL, 0x21f8d4a76f74541aULL}},
        {{0xcaf5185423a7d23aULL, 0x7cef6acb145b6413ULL, 0x2879dd439b019b8bULL, 0x71449cdeca4f0007ULL,
          0xdebaf4a2c5534527ULL, 0xa1b4e791d1b86560ULL, 0x1e0f563c601bb8dcULL}},
        {{0x274315837455b919ULL, 0x82039e4221ff3507ULL, 0x00346cebad16a036ULL, 0x0177bfd6654e681eULL,
          0xddff621b5db3f897ULL, 0x0cc61570301497a7ULL, 0x115ea2305a78f646ULL}},
        {{0x392236e9cf2976c2ULL, 0xd8ab17c84b9f03cdULL, 0x8a8e6755f9d82fd1ULL, 0x7532834528cd5a64ULL,
          0x0b0bcc3fb6f2161cULL, 0x76a2ffcb7d47679dULL, 0x25ed2192b203c1feULL}}};
    Aff<Fp2Ops> g;
    gmh::Fp t;
    gmh::to_mont(t, c[0]);
    gmh::to_limbs28(g.x.c0.v, t);
    gmh::to_mont(t, c[1]);
    gmh::to_limbs28(g.x.c1.v, t);
    gmh::to_mont(t, c[2]);
    gmh::to_limbs28(g.y.c0.v, t);
    gmh::to_mont(t, c[3]);
    gmh::to_limbs28(g.y.c1.v, t);
    return g;
}

template <class O>
int bases_create(gm_bases** out, const void* host_points, const uint64_t* host_k, int n, int montgomery,
                 uint32_t gpu) {
    using Lay = gm::AbiLayout<O>;
    if (!out || n <= 0) return -1;
    *out = nullptr;
    DevCtx* c = get_ctx(gpu);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->dev) != hipSuccess) return -2;
    void* d = nullptr;
    if (hipMalloc(&d, (size_t)n * sizeof(Aff<O>)) != hipSuccess) return -3;
    void* tmp = nullptr;
    const size_t tbytes = host_points ? (size_t)n * Lay::AFF_BYTES : (size_t)n * 8;
    if (hipMalloc(&tmp, tbytes) != hipSuccess) {
        hipFree(d);
        return -3;
    }
    hipError_t e = h2d_staged(c, tmp, host_points ? host_points : (const void*)host_k, tbytes);
    if (e == hipSuccess) {
        if (host_points)
            e = gm::Engine<O>::import_aff(c->stream, reinterpret_cast<Aff<O>*>(d), tmp, n, montgomery == 0);
        else
            e = gm::Engine<O>::gen_bases(c->stream, reinterpret_cast<Aff<O>*>(d), reinterpret_cast<uint64_t*>(tmp), n,
                                         generator_dev<O>());
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    if (e != hipSuccess) {
        hipFree(d);
        return -4;
    }
    gm_bases* b = new gm_bases{HostOf<O>::CURVE, n, c->dev, d};
    *out = b;
    return 0;
}

template <class O>
int resident_batch(void* results_affine, const gm_bases* b, const gm_scalars* s, const gm_msm_spec* specs, int count,
                   int bits, int device_combine) {
    using HF = typename HostOf<O>::HF;
    if (!results_affine || !b || !s || !specs || count <= 0 || b->dev != s->dev) return -1;
    std::vector<gm::MsmSpec> sp((size_t)count);
    for (int j = 0; j < count; ++j) {
        const gm_msm_spec& q = specs[j];
        if (q.n <= 0 || q.base_offset < 0 || q.scalar_offset < 0) return -1;
        if ((long)q.base_offset + q.n > b->n || (long)q.scalar_offset + q.n > s->n) return -1;
        sp[j] = {q.n, q.base_offset, q.scalar_offset, q.window, q.seg};
    }
    DevCtx* c = get_ctx((uint32_t)b->dev);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->dev) != hipSuccess) return -2;
    std::vector<AffResult<HF>> res;
    int rc = run_batch_locked<O>(c, reinterpret_cast<const Aff<O>*>(b->d), s->d, s->limbs, bits, sp.data(), count,
                                 device_combine != 0, res);
    if (rc) return rc;
    const size_t rec = HostOf<O>::CURVE == gm::CURVE_G1 ? 120 : 232;
    for (int j = 0; j < count; ++j)
        write_affine_abi<HF>(reinterpret_cast<uint8_t*>(results_affine) + (size_t)j * rec, res[j].finite, res[j].x,
                             res[j].y);
    return 0;
}

// ---------------------------------------------------------- batch objects

}  // namespace

// A fixed batch of MSMs over resident data with its own stream, workspace and
// (optionally) a captured hipGraph of the whole pipeline, replayed per run.
struct gm_batch {
    int curve = 0, dev = 0, count = 0;
    std::vector<gm::MsmDesc> desc;
    Plan p{};
    const gm_bases* b = nullptr;
    const gm_scalars* s = nullptr;
    uint8_t* ws = nullptr;
    uint8_t* pinned = nullptr;
    size_t res_bytes = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = false;
    bool device_combine = false;
    bool in_flight = false;
};

namespace {

template <class O>
hipError_t batch_enqueue(gm_batch* bt) {
    GM_TRY_HOST(gm::Engine<O>::run(bt->stream, bt->p, reinterpret_cast<const Aff<O>*>(bt->b->d), bt->s->d, bt->ws,
                                   nullptr));
    // pinned: the error words, then the results ERR_BYTES on
    if (bt->device_combine) {
        GM_TRY_HOST(gm::Engine<O>::device_combine(bt->stream, bt->p, bt->ws));
        GM_TRY_HOST(hipMemcpyAsync(bt->pinned + gm::ERR_BYTES, bt->ws + bt->p.o_out, bt->res_bytes,
                                   hipMemcpyDeviceToHost, bt->stream));
        GM_TRY_HOST(hipMemcpyAsync(bt->pinned, bt->ws + bt->p.o_err, 16, hipMemcpyDeviceToHost, bt->stream));
    } else {
        // error words and window sums in one copy
        GM_TRY_HOST(hipMemcpyAsync(bt->pinned, bt->ws + bt->p.o_err, gm::ERR_BYTES + bt->res_bytes,
                                   hipMemcpyDeviceToHost, bt->stream));
    }
    return hipSuccess;
}

hipError_t batch_enqueue_any(gm_batch* bt) {
    return bt->curve == GM_CURVE_G1 ? batch_enqueue<FpOps>(bt) : batch_enqueue<Fp2Ops>(bt);
}

void batch_release(gm_batch* bt) {
    if (!bt) return;
    hipSetDevice(bt->dev);
    if (bt->stream) hipStreamSynchronize(bt->stream);
    if (bt->exec) hipGraphExecDestroy(bt->exec);
    if (bt->graph) hipGraphDestroy(bt->graph);
    if (bt->stream) hipStreamDestroy(bt->stream);
    if (bt->ws) hipFree(bt->ws);
    if (bt->pinned) hipHostFree(bt->pinned);
    delete bt;
}

// Host window combines of many MSMs, spread over host threads.
template <class HF>
void host_combine_many(std::vector<AffResult<HF>>& out, const uint32_t* raw, const gm::MsmDesc* desc, int count) {
    constexpr int W = words_of<HF>();
    out.resize((size_t)count);
    auto work = [&](int j0, int j1) {
        for (int j = j0; j < j1; ++j)
            out[j].finite = host_combine<HF>(out[j].x, out[j].y, raw + (size_t)desc[j].win_base * 4 * W,
                                             (int)desc[j].nw, (int)desc[j].c);
    };
    const int nt = count < 8 ? 1 : (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (nt == 1) {
        work(0, count);
        return;
    }
    std::vector<std::thread> th;
    const int per = (count + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int j0 = t * per, j1 = std::min(count, j0 + per);
        if (j0 < j1) th.emplace_back(work, j0, j1);
    }
    for (auto& t : th) t.join();
}

template <class O>
int batch_results(gm_batch* bt, void* results_affine) {
    using HF = typename HostOf<O>::HF;
    uint32_t err;
    memcpy(&err, bt->pinned, 4);
    if (err) return -5;
    const uint8_t* res_h = bt->pinned + gm::ERR_BYTES;
    constexpr int W = words_of<HF>();
    const size_t rec = HostOf<O>::CURVE == gm::CURVE_G1 ? 120 : 232;
    std::vector<AffResult<HF>> res((size_t)bt->count);
    if (bt->device_combine) {
        for (int j = 0; j < bt->count; ++j) {
            const uint32_t* a = reinterpret_cast<const uint32_t*>(res_h + (size_t)j * sizeof(Aff<O>));
            res[j].finite = a[15] != gm::INF_MARK;
            if (res[j].finite) {
                dev_elem_to_host(res[j].x, a);
                dev_elem_to_host(res[j].y, a + W);
            }
        }
    } else {
        host_combine_many<HF>(res, reinterpret_cast<const uint32_t*>(res_h), bt->desc.data(), bt->count);
    }
    for (int j = 0; j < bt->count; ++j)
        write_affine_abi<HF>(reinterpret_cast<uint8_t*>(results_affine) + (size_t)j * rec, res[j].finite, res[j].x,
                             res[j].y);
    return 0;
}

}  // namespace

// ------------------------------------------------------- point formats

namespace {

// One format / validation pass over n host records: upload, launch, download.
// Returns the number of records with a non-zero status, or a negative code.
template <class O>
int fmt_run(int op, const void* host_in, size_t in_bytes, void* host_out, size_t out_bytes, int n, int arg,
            int32_t* status, uint32_t gpu) {
    DevCtx* c = get_ctx(gpu);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->dev) != hipSuccess) return -2;
    const size_t o_out = gm::al_size(in_bytes), o_st = o_out + gm::al_size(out_bytes);
    if (ensure_ws(c, o_st + gm::al_size((size_t)n * 4)) != hipSuccess) return -3;
    uint8_t* d = c->ws;
    int32_t* d_st = reinterpret_cast<int32_t*>(d + o_st);
    if (hipMemcpyAsync(d, host_in, in_bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) return -4;
    if (gm::launch_fmt<O>(c->stream, op, d, d + o_out, n, arg, d_st) != hipSuccess) return -4;
    if (out_bytes && hipMemcpyAsync(host_out, d + o_out, out_bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return -4;
    std::vector<int32_t> st(status ? (size_t)n : 0);
    if (status && hipMemcpyAsync(st.data(), d_st, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return -4;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -4;
    int bad = 0;
    for (int i = 0; i < (int)st.size(); ++i) {
        status[i] = st[i];
        bad += st[i] != 0;
    }
    return bad;
}

int fmt_curve(int curve, int op, const void* in, size_t in_rec, void* out, size_t out_rec, int n, int arg,
              int32_t* status, uint32_t gpu) {
    if (n < 0 || (n > 0 && !in)) return -1;
    if (n == 0) return 0;
    if (curve == GM_CURVE_G1)
        return fmt_run<FpOps>(op, in, in_rec * n, out, out_rec * n, n, arg, status, gpu);
    if (curve == GM_CURVE_G2)
        return fmt_run<Fp2Ops>(op, in, in_rec * n, out, out_rec * n, n, arg, status, gpu);
    return -1;
}

size_t aff_rec(int curve) { return curve == GM_CURVE_G1 ? sizeof(G1Point) : sizeof(G2Point); }

}  // namespace

// ================================================================= C ABI

extern "C" {

// reference rust-api/src/c_wrapper.cu:12-19
void affine_to_projective_g1_wrapper(G1ProjectivePoint* proj, const G1Point* affine) {
    memset(proj, 0, sizeof *proj);
    if (affine->infinity) return;
    proj->X = affine->x;
    proj->Y = affine->y;
    memcpy(proj->Z.limb, gmh::ONE.l, 56);
}
void affine_to_projective_g2_wrapper(G2ProjectivePoint* proj, const G2Point* affine) {
    memset(proj, 0, sizeof *proj);
    if (affine->infinity) return;
    proj->X = affine->x;
    proj->Y = affine->y;
    memcpy(proj->Z.c0.limb, gmh::ONE.l, 56);
}

// reference rust-api/src/c_wrapper.cu:22-29 (src/curve.cu:754-778)
void projective_to_affine_g1_wrapper(G1Point* affine, const G1ProjectivePoint* proj) {
    gmh::Fp X, Y, Z, zi, x, y;
    get_elem(X, proj->X.limb);
    get_elem(Y, proj->Y.limb);
    get_elem(Z, proj->Z.limb);
    G1Point r;
    memset(&r, 0, sizeof r);
    if (gmh::is_zero(Z)) {
        r.infinity = true;
    } else {
        gmh::inv(zi, Z);
        gmh::mul(x, X, zi);
        gmh::mul(y, Y, zi);
        memcpy(r.x.limb, x.l, 56);
        memcpy(r.y.limb, y.l, 56);
    }
    *affine = r;
}
void projective_to_affine_g2_wrapper(G2Point* affine, const G2ProjectivePoint* proj) {
    gmh::Fp2 X, Y, Z, zi, x, y;
    get_elem(X, proj->X.c0.limb);
    get_elem(Y, proj->Y.c0.limb);
    get_elem(Z, proj->Z.c0.limb);
    G2Point r;
    memset(&r, 0, sizeof r);
    if (gmh::is_zero(Z)) {
        r.infinity = true;
    } else {
        gmh::inv(zi, Z);
        gmh::mul(x, X, zi);
        gmh::mul(y, Y, zi);
        put_elem(r.x.c0.limb, x);
        put_elem(r.y.c0.limb, y);
    }
    *affine = r;
}

// reference rust-api/src/c_wrapper.cu:32-59
void g1_point_at_infinity_wrapper(G1Point* point) {
    memset(point, 0, sizeof *point);
    point->infinity = true;
}
void g2_point_at_infinity_wrapper(G2Point* point) {
    memset(point, 0, sizeof *point);
    point->infinity = true;
}
void g1_projective_point_at_infinity_wrapper(G1ProjectivePoint* point) { memset(point, 0, sizeof *point); }
void g2_projective_point_at_infinity_wrapper(G2ProjectivePoint* point) { memset(point, 0, sizeof *point); }
bool g1_is_infinity_wrapper(const G1Point* point) { return point->infinity; }
bool g2_is_infinity_wrapper(const G2Point* point) { return point->infinity; }

// reference rust-api/src/c_wrapper.cu:64-172
int g1_msm_wrapper(G1ProjectivePoint* result, const G1Point* points, const uint64_t* scalars, int n,
                   uint32_t gpu_index) {
    return msm_host_ptrs<FpOps>(result, points, scalars, 1, 64, n, gpu_index, true);
}

// reference rust-api/src/c_wrapper.cu:175-258
int g2_msm_wrapper(G2ProjectivePoint* result, const G2Point* points, const uint64_t* scalars, int n,
                   uint32_t gpu_index) {
    return msm_host_ptrs<Fp2Ops>(result, points, scalars, 1, 64, n, gpu_index, false);
}

// reference rust-api/src/ffi.rs:67 (src/curve.cu:833-838)
void init_device_generators(const void* stream, uint32_t gpu_index) {
    (void)stream;
    get_ctx(gpu_index);
}

// reference rust-api/src/c_wrapper.cu:261-283
void g1_from_montgomery_wrapper(G1Point* result, const G1Point* point) {
    G1Point r;
    memset(&r, 0, sizeof r);
    if (point->infinity) {
        r.infinity = true;
    } else {
        gmh::Fp a, b;
        get_elem(a, point->x.limb);
        gmh::from_mont(b, a);
        memcpy(r.x.limb, b.l, 56);
        get_elem(a, point->y.limb);
        gmh::from_mont(b, a);
        memcpy(r.y.limb, b.l, 56);
    }
    *result = r;
}
void g2_from_montgomery_wrapper(G2Point* result, const G2Point* point) {
    G2Point r;
    memset(&r, 0, sizeof r);
    if (point->infinity) {
        r.infinity = true;
    } else {
        const Fp* in[4] = {&point->x.c0, &point->x.c1, &point->y.c0, &point->y.c1};
        Fp* out[4] = {&r.x.c0, &r.x.c1, &r.y.c0, &r.y.c1};
        for (int k = 0; k < 4; ++k) {
            gmh::Fp a, b;
            get_elem(a, in[k]->limb);
            gmh::from_mont(b, a);
            memcpy(out[k]->limb, b.l, 56);
        }
    }
    *result = r;
}
// reference rust-api/src/c_wrapper.cu:286-288
void fp_to_montgomery_wrapper(Fp* result, const Fp* value) {
    gmh::Fp a, b;
    get_elem(a, value->limb);
    gmh::to_mont(b, a);
    memcpy(result->limb, b.l, 56);
}

// ------------------------------------------------------------- extensions

int g1_msm_multi_wrapper(G1ProjectivePoint* result, const G1Point* points, const uint64_t* scalars,
                         int scalar_limbs, int n, uint32_t gpu_index, uint32_t flags) {
    const int bits = (flags & GM_SCALARS_LT_R) ? 299 : 64 * scalar_limbs;
    return msm_host_ptrs<FpOps>(result, points, scalars, scalar_limbs, bits, n, gpu_index,
                                (flags & GM_POINTS_MONTGOMERY) == 0);
}
int g2_msm_multi_wrapper(G2ProjectivePoint* result, const G2Point* points, const uint64_t* scalars,
                         int scalar_limbs, int n, uint32_t gpu_index, uint32_t flags) {
    const int bits = (flags & GM_SCALARS_LT_R) ? 299 : 64 * scalar_limbs;
    return msm_host_ptrs<Fp2Ops>(result, points, scalars, scalar_limbs, bits, n, gpu_index,
                                 (flags & GM_POINTS_MONTGOMERY) == 0);
}

int gm_plan_query(int curve, int n, int scalar_limbs, int scalar_bits, int window, int* c_out, int* windows_out,
                  size_t* workspace_bytes_out) {
    if (curve != GM_CURVE_G1 && curve != GM_CURVE_G2) return -1;
    gm::MsmSpec spec = {n, 0, 0, window, 0};
    gm::MsmDesc d;
    Plan p;
    if (!gm::make_plan(p, &d, curve, scalar_limbs, scalar_bits, &spec, 1)) return -1;
    if (c_out) *c_out = (int)d.c;
    if (windows_out) *windows_out = (int)d.nw;
    if (workspace_bytes_out) *workspace_bytes_out = p.total;
    return 0;
}

int gm_bases_upload(gm_bases** out, int curve, const void* host_points, int n, int montgomery, uint32_t gpu_index) {
    if (!host_points) return -1;
    if (curve == GM_CURVE_G1) return bases_create<FpOps>(out, host_points, nullptr, n, montgomery, gpu_index);
    if (curve == GM_CURVE_G2) return bases_create<Fp2Ops>(out, host_points, nullptr, n, montgomery, gpu_index);
    return -1;
}

int gm_bases_from_k(gm_bases** out, int curve, const uint64_t* host_k, int n, uint32_t gpu_index) {
    if (!host_k) return -1;
    if (curve == GM_CURVE_G1) return bases_create<FpOps>(out, nullptr, host_k, n, 1, gpu_index);
    if (curve == GM_CURVE_G2) return bases_create<Fp2Ops>(out, nullptr, host_k, n, 1, gpu_index);
    return -1;
}

int gm_bases_download(const gm_bases* b, void* host_points, int offset, int n) {
    if (!b || !host_points || offset < 0 || n <= 0 || (long)offset + n > b->n) return -1;
    DevCtx* c = get_ctx((uint32_t)b->dev);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->dev) != hipSuccess) return -2;
    const size_t rec = b->curve == GM_CURVE_G1 ? 120 : 232;
    void* tmp = nullptr;
    if (hipMalloc(&tmp, (size_t)n * rec) != hipSuccess) return -3;
    hipError_t e;
    if (b->curve == GM_CURVE_G1)
        e = gm::Engine<FpOps>::export_aff(c->stream, tmp, reinterpret_cast<const Aff<FpOps>*>(b->d) + offset, n);
    else
        e = gm::Engine<Fp2Ops>::export_aff(c->stream, tmp, reinterpret_cast<const Aff<Fp2Ops>*>(b->d) + offset, n);
    if (e == hipSuccess) e = hipMemcpyAsync(host_points, tmp, (size_t)n * rec, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(tmp);
    return e == hipSuccess ? 0 : -4;
}

int gm_bases_count(const gm_bases* b) { return b ? b->n : 0; }

void gm_bases_free(gm_bases* b) {
    if (!b) return;
    hipSetDevice(b->dev);
    hipFree(b->d);
    delete b;
}

int gm_scalars_upload(gm_scalars** out, const uint64_t* host, int n, int limbs, uint32_t gpu_index) {
    if (!out || !host || n <= 0 || limbs < 1 || limbs > 8) return -1;
    *out = nullptr;
    DevCtx* c = get_ctx(gpu_index);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->dev) != hipSuccess) return -2;
    uint64_t* d = nullptr;
    if (hipMalloc(&d, (size_t)n * limbs * 8) != hipSuccess) return -3;
    if (h2d_staged(c, d, host, (size_t)n * limbs * 8) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        hipFree(d);
        return -4;
    }
    *out = new gm_scalars{n, limbs, c->dev, d};
    return 0;
}

void gm_scalars_free(gm_scalars* s) {
    if (!s) return;
    hipSetDevice(s->dev);
    hipFree(s->d);
    delete s;
}

int gm_msm_resident(void* result_affine, const gm_bases* b, int base_offset, const gm_scalars* s, int scalar_offset,
                    int n, int scalar_bits, int window, int device_combine) {
    gm_msm_spec spec = {n, base_offset, scalar_offset, window, 0};
    return gm_msm_batch(result_affine, b, s, &spec, 1, scalar_bits, device_combine);
}

int gm_msm_batch(void* results_affine, const gm_bases* b, const gm_scalars* s, const gm_msm_spec* specs, int count,
                 int scalar_bits, int device_combine) {
    if (!b) return -1;
    if (b->curve == GM_CURVE_G1)
        return resident_batch<FpOps>(results_affine, b, s, specs, count, scalar_bits, device_combine);
    return resident_batch<Fp2Ops>(results_affine, b, s, specs, count, scalar_bits, device_combine);
}

int gm_batch_create(gm_batch** out, const gm_bases* b, const gm_scalars* s, const gm_msm_spec* specs, int count,
                    int scalar_bits, int flags) {
    const int use_graph = (flags & GM_BATCH_GRAPH) != 0;
    if (!out || !b || !s || !specs || count <= 0 || b->dev != s->dev) return -1;
    *out = nullptr;
    std::vector<gm::MsmSpec> sp((size_t)count);
    for (int j = 0; j < count; ++j) {
        const gm_msm_spec& q = specs[j];
        if (q.n <= 0 || q.base_offset < 0 || q.scalar_offset < 0) return -1;
        if ((long)q.base_offset + q.n > b->n || (long)q.scalar_offset + q.n > s->n) return -1;
        sp[j] = {q.n, q.base_offset, q.scalar_offset, q.window, q.seg};
    }
    if (hipSetDevice(b->dev) != hipSuccess) return -2;
    gm_batch* bt = new gm_batch();
    bt->curve = b->curve;
    bt->dev = b->dev;
    bt->count = count;
    bt->b = b;
    bt->s = s;
    bt->desc.resize((size_t)count);
    if (!gm::make_plan(bt->p, bt->desc.data(), b->curve, s->limbs, scalar_bits, sp.data(), count)) {
        delete bt;
        return -1;
    }
    bt->device_combine = (flags & GM_BATCH_DEVICE_COMBINE) != 0;
    bt->res_bytes = bt->device_combine ? (size_t)count * gm::aff_bytes(b->curve)
                                       : (size_t)bt->p.win_total * gm::xyzz_bytes(b->curve);
    if (hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking) != hipSuccess) {
        batch_release(bt);
        return -2;
    }
    if (hipMalloc(&bt->ws, bt->p.total) != hipSuccess ||
        hipHostMalloc((void**)&bt->pinned, gm::ERR_BYTES + gm::al_size(bt->res_bytes), hipHostMallocDefault) !=
            hipSuccess) {
        batch_release(bt);
        return -3;
    }
    if (hipMemcpy(bt->ws + bt->p.o_desc, bt->desc.data(), (size_t)count * sizeof(gm::MsmDesc),
                  hipMemcpyHostToDevice) != hipSuccess) {
        batch_release(bt);
        return -4;
    }
    if (use_graph) {
        hipError_t e = hipStreamBeginCapture(bt->stream, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            hipError_t e2 = batch_enqueue_any(bt);
            e = hipStreamEndCapture(bt->stream, &bt->graph);
            if (e2 != hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            batch_release(bt);
            return -4;
        }
        bt->use_graph = true;
    }
    *out = bt;
    return 0;
}

int gm_batch_launch(gm_batch* bt) {
    if (!bt || bt->in_flight) return -1;
    if (hipSetDevice(bt->dev) != hipSuccess) return -2;
    hipError_t e = bt->use_graph ? hipGraphLaunch(bt->exec, bt->stream) : batch_enqueue_any(bt);
    if (e != hipSuccess) return -4;
    bt->in_flight = true;
    return 0;
}

int gm_batch_wait(gm_batch* bt, void* results_affine) {
    if (!bt || !bt->in_flight || !results_affine) return -1;
    if (hipSetDevice(bt->dev) != hipSuccess) return -2;
    hipError_t e = hipStreamSynchronize(bt->stream);
    bt->in_flight = false;
    if (e != hipSuccess) return -4;
    return bt->curve == GM_CURVE_G1 ? batch_results<FpOps>(bt, results_affine)
                                    : batch_results<Fp2Ops>(bt, results_affine);
}

int gm_batch_run(gm_batch* bt, void* results_affine) {
    int rc = gm_batch_launch(bt);
    if (rc) return rc;
    return gm_batch_wait(bt, results_affine);
}

int gm_batch_count(const gm_batch* bt) { return bt ? bt->count : 0; }

size_t gm_batch_workspace_bytes(const gm_batch* bt) { return bt ? bt->p.total : 0; }

void gm_batch_free(gm_batch* bt) { batch_release(bt); }

int gm_set_profiling(uint32_t gpu_index, int on) {
    DevCtx* c = get_ctx(gpu_index);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    c->prof.on = on != 0;
    c->prof_valid = false;
    return 0;
}

int gm_last_msm_info(uint32_t gpu_index, int* window, int* windows, long long* entries) {
    DevCtx* c = get_ctx(gpu_index);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (window) *window = c->last_c;
    if (windows) *windows = c->last_nw;
    if (entries) *entries = c->last_entries;
    return 0;
}

int gm_profile_last(uint32_t gpu_index, float* stage_ms, int count) {
    DevCtx* c = get_ctx(gpu_index);
    if (!c) return -2;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->prof_valid || !stage_ms) return -1;
    for (int i = 0; i < count && i + 1 < gm::ST_COUNT; ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, c->prof.ev[i], c->prof.ev[i + 1]) != hipSuccess) return -4;
        stage_ms[i] = ms;
    }
    return 0;
}

int gm_point_encoded_size(int curve, int compressed) {
    const int e = curve == GM_CURVE_G1 ? 56 : curve == GM_CURVE_G2 ? 112 : 0;
    return compressed ? e : 2 * e;
}

int gm_points_serialize(int curve, void* out_bytes, const void* points, int n, int compressed, uint32_t gpu_index) {
    const int sz = gm_point_encoded_size(curve, compressed);
    if (sz == 0 || (n > 0 && !out_bytes)) return -1;
    return fmt_curve(curve, gm::FMT_OP_SERIALIZE, points, aff_rec(curve), out_bytes, (size_t)sz, n,
                     compressed ? 1 : 0, nullptr, gpu_index);
}

int gm_points_deserialize(int curve, void* points, const void* in_bytes, int n, int compressed, int checks,
                          int32_t* status, uint32_t gpu_index) {
    const int sz = gm_point_encoded_size(curve, compressed);
    if (sz == 0 || (n > 0 && (!points || !status)) || (checks & ~3)) return -1;
    return fmt_curve(curve, compressed ? gm::FMT_OP_DESERIALIZE_C : gm::FMT_OP_DESERIALIZE_U, in_bytes, (size_t)sz,
                     points, aff_rec(curve), n, checks, status, gpu_index);
}

int gm_points_check(int curve, const void* points, int n, int checks, int32_t* status, uint32_t gpu_index) {
    if ((n > 0 && !status) || (checks & ~3)) return -1;
    return fmt_curve(curve, gm::FMT_OP_CHECK, points, aff_rec(curve), nullptr, 0, n, checks, status, gpu_index);
}

int gm_points_to_canonical(int curve, void* out, const void* points, int n, uint32_t gpu_index) {
    if (n > 0 && !out) return -1;
    return fmt_curve(curve, gm::FMT_OP_CANONICAL, points, aff_rec(curve), out, aff_rec(curve), n, 1, nullptr,
                     gpu_index);
}

int gm_points_from_canonical(int curve, void* out, const void* points, int n, uint32_t gpu_index) {
    if (n > 0 && !out) return -1;
    return fmt_curve(curve, gm::FMT_OP_CANONICAL, points, aff_rec(curve), out, aff_rec(curve), n, 0, nullptr,
                     gpu_index);
}

}  // extern "C"
