// Pinned staging for host -> device copies of caller (pageable) buffers, used
// by the per-device context of the C ABI and by each slot of the request
// server. One pair of 4 MiB pinned slots per worker thread, each slot with the
// event of its last DMA.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace gmc {

struct Staging {
    static constexpr int WORKERS = 8;
    static constexpr int SLOTS = 2;
    static constexpr size_t SLOT_BYTES = (size_t)4 << 20;
    // copies below this size go straight to hipMemcpyAsync
    static constexpr size_t MIN_BYTES = (size_t)8 << 20;
    uint8_t* buf = nullptr;  // WORKERS * SLOTS * SLOT_BYTES
    hipEvent_t ev[WORKERS][SLOTS] = {};
    bool failed = false;  // allocation failed once: plain copies from then on
};

inline void staging_release(Staging& st) {
    for (auto& row : st.ev)
        for (hipEvent_t& e : row)
            if (e) {
                hipEventDestroy(e);
                e = nullptr;
            }
    if (st.buf) hipHostFree(st.buf);
    st.buf = nullptr;
}

inline bool staging_ensure(Staging& st) {
    if (st.buf) return true;
    if (st.failed) return false;
    const size_t bytes = (size_t)Staging::WORKERS * Staging::SLOTS * Staging::SLOT_BYTES;
    if (hipHostMalloc((void**)&st.buf, bytes, hipHostMallocDefault) != hipSuccess) {
        st.buf = nullptr;
        st.failed = true;
        return false;
    }
    for (int w = 0; w < Staging::WORKERS; ++w)
        for (int k = 0; k < Staging::SLOTS; ++k)
            if (hipEventCreateWithFlags(&st.ev[w][k], hipEventDisableTiming) != hipSuccess) {
                st.ev[w][k] = nullptr;
                staging_release(st);
                st.failed = true;
                return false;
            }
    return true;
}

// One host -> device copy of a pageable caller buffer.
struct H2D {
    void* dst;
    const void* src;
    size_t bytes;
};

// Host -> device copies of pageable caller buffers on `stream` of device
// `dev`. Large copies are split into one contiguous part of their
// concatenation per worker thread; each worker copies its part slot by slot
// into pinned memory (several host threads together move pageable memory
// faster than one) and queues each slot's DMA on the stream, waiting on the
// slot's previous DMA before refilling it, so host copies and DMAs overlap.
// Every DMA is queued before this returns, so work queued on the stream
// afterwards sees the data; the caller's buffers are no longer read. One
// caller at a time per Staging.
inline hipError_t h2d_staged(Staging& st, int dev, hipStream_t stream, const H2D* cp, int ncp) {
    size_t total = 0;
    for (int i = 0; i < ncp; ++i) total += cp[i].bytes;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int workers = (int)std::min<size_t>(std::min<unsigned>(Staging::WORKERS, hw), total / Staging::SLOT_BYTES);
    if (total < Staging::MIN_BYTES || workers < 2 || !staging_ensure(st)) {
        for (int i = 0; i < ncp; ++i)
            if (cp[i].bytes) {
                const hipError_t e = hipMemcpyAsync(cp[i].dst, cp[i].src, cp[i].bytes, hipMemcpyHostToDevice, stream);
                if (e != hipSuccess) return e;
            }
        return hipSuccess;
    }
    const size_t part = (total + workers - 1) / workers;
    std::vector<hipError_t> err(workers, hipSuccess);
    auto run = [&](int w) {
        hipError_t e = hipSetDevice(dev);
        const size_t a = std::min(total, (size_t)w * part), b = std::min(total, a + part);
        int k = 0;
        size_t base = 0;  // offset of copy i in the concatenation
        for (int i = 0; i < ncp && e == hipSuccess; base += cp[i].bytes, ++i) {
            const size_t lo = std::max(a, base), hi = std::min(b, base + cp[i].bytes);
            for (size_t o = lo; o < hi && e == hipSuccess; o += Staging::SLOT_BYTES, k ^= 1) {
                const size_t len = std::min(Staging::SLOT_BYTES, hi - o);
                uint8_t* slot = st.buf + ((size_t)w * Staging::SLOTS + k) * Staging::SLOT_BYTES;
                e = hipEventSynchronize(st.ev[w][k]);  // the slot's previous DMA is done
                if (e != hipSuccess) break;
                memcpy(slot, (const uint8_t*)cp[i].src + (o - base), len);
                e = hipMemcpyAsync((uint8_t*)cp[i].dst + (o - base), slot, len, hipMemcpyHostToDevice, stream);
                if (e == hipSuccess) e = hipEventRecord(st.ev[w][k], stream);
            }
        }
        err[w] = e;
    };
    std::vector<std::thread> th;
    int started = 1;
    try {
        th.reserve(workers - 1);
        for (; started < workers; ++started) th.emplace_back(run, started);
    } catch (...) {
        // a thread that cannot start leaves its part (and the later ones) to this thread
    }
    for (int w = started; w < workers; ++w) run(w);
    run(0);
    for (auto& t : th) t.join();
    for (hipError_t e : err)
        if (e != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace gmc
