// Served MSM request loop: continuous batching of mixed G1/G2 requests over
// bases resident in HBM.
//
// Requests (curve, base offset, n, scalars) are queued by gm_server_submit.
// A scheduler thread forms a batch as soon as a stream slot is free, admitting
// requests while the batch's device footprint (pipeline workspace + scalars)
// fits what is left of the HBM budget, then launches the whole pipeline on
// that slot's stream. Several slots run at once, so batches that form while
// others are on the GPU overlap on the device.
//
// Requests of at most small_max_n points form their own batches and go first
// (FIFO among them), and one slot is kept for them: batches of larger
// requests occupy at most slots - 1 slots, so a small request never queues
// behind large MSMs (a 2^20 G2 MSM holds its slot for some 16 ms). Every
// batch holds one curve, scalar width and size class.
//
// The scheduler only forms batches. A worker thread per slot stages the
// batch's scalars through the slot's pinned buffers (several host threads;
// the caller's memory is pageable), queues the pipeline on the slot's stream,
// waits for it and releases the slot as soon as the device is done; the window
// combines then run on a pool of host threads, which publish each request's
// affine result and latency (submit -> result ready). Batches complete in
// whatever order they finish, not in launch order, and a slot takes its next
// batch while the host still combines the last one.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gpu_msm_ext.h"
#include "engine.hpp"
#include "host_combine.hpp"
#include "host_field.hpp"
#include "resident.hpp"
#include "staging.hpp"

#define GM_TRY_S(x)                      \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

namespace {

using gm::Aff;
using gm::Fp2Ops;
using gm::FpOps;
using gm::Xyzz;

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t ticket = 0;
    int curve = 0, base_off = 0, n = 0, limbs = 0, bits = 0;
    std::vector<uint64_t> scalars;
    const uint64_t* borrowed = nullptr;  // caller-owned scalars (GM_SUBMIT_BORROW)
    double t_submit = 0, t_done = 0;
    bool done = false;
    int rc = 0;
    uint8_t result[232];
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint8_t* ws = nullptr;
    size_t ws_cap = 0;
    uint8_t* d_sc = nullptr;
    size_t sc_cap = 0;
    uint8_t* pinned = nullptr;
    size_t pinned_cap = 0;
    gmc::Staging stage;  // pinned staging of the batch's scalars
    bool busy = false;
    bool launched = false;  // device work queued by the slot's worker
    bool large = false;     // a batch of the large size class
    bool fail = false;
    std::vector<Request*> reqs;
    std::vector<gm::MsmDesc> desc;
    gm::Plan p{};
    int curve = 0;
    size_t footprint = 0;
    size_t res_bytes = 0;
};

// One request's window combine, queued once its batch has left the device.
struct CombineJob {
    Request* r = nullptr;
    int curve = 0, nw = 0, c = 0;
    std::vector<uint32_t> raw;  // nw XYZZ window sums, device words
};

// ---- planner MLP (host inference; same network as csrc/mlp.hip)

struct HostMlp {
    std::vector<float> P;
    bool ok() const { return P.size() == GM_MLP_PARAMS; }
    // (c, seg) for one MSM of scalars below 2^bits, or (0, 0) to use the
    // heuristic
    void plan(int curve, int n, int batch, int bits, int& c, int& seg) const {
        c = 0;
        seg = 0;
        if (!ok()) return;
        const float x[3] = {(float)((log2((double)std::max(n, 1)) - 13.0) / 5.0),
                            (float)((log2((double)std::max(batch, 1)) - 3.0) / 3.0), (float)(curve - 1)};
        float h1[32], h2[32], y[2];
        const float *W1 = P.data(), *b1 = W1 + 96, *W2 = b1 + 32, *b2 = W2 + 1024, *W3 = b2 + 32, *b3 = W3 + 64;
        for (int j = 0; j < 32; ++j) {
            float a = b1[j];
            for (int i = 0; i < 3; ++i) a = fmaf(W1[j * 3 + i], x[i], a);
            h1[j] = a > 0.f ? a : 0.f;
        }
        for (int j = 0; j < 32; ++j) {
            float a = b2[j];
            for (int i = 0; i < 32; ++i) a = fmaf(W2[j * 32 + i], h1[i], a);
            h2[j] = a > 0.f ? a : 0.f;
        }
        for (int k = 0; k < 2; ++k) {
            float a = b3[k];
            for (int i = 0; i < 32; ++i) a = fmaf(W3[k * 32 + i], h2[i], a);
            y[k] = a;
        }
        // outputs: a correction to the heuristic's width, and log2 seg
        const double dc = y[0] * 2.0, ls = y[1] * 1.5 + 3.0;
        if (!std::isfinite(dc) || !std::isfinite(ls) || std::fabs(dc) > 20.0) return;  // heuristic fallback
        // c inside the sweeps' candidate band around the heuristic, at the
        // nearest width whose top window is not sparse
        const int h = gm::choose_window(curve, std::max(n, 1), bits);
        const int lo = std::max(3, h - 3), hi = std::min(20, h + 2);
        const double ct = std::min((double)hi, std::max((double)lo, h + dc));
        c = h;
        for (int k = lo; k <= hi; ++k)
            if (gm::window_ok(bits, k) && std::fabs(k - ct) < std::fabs(c - ct)) c = k;
        seg = 1 << (int)lround(std::min(5.0, std::max(2.0, ls)));
        seg = std::min(seg, 1 << (c - 1));
    }
};

}  // namespace

struct gm_server {
    gm_server_config cfg{};
    const gm_bases* pool[3] = {nullptr, nullptr, nullptr};
    HostMlp mlp;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_slot, cv_job;
    std::deque<Request*> queue;
    std::deque<CombineJob> jobs;
    int combining = 0;  // jobs taken by a combine thread and not yet published
    std::unordered_map<uint64_t, Request*> table;
    std::vector<Slot> slots;
    size_t inflight_bytes = 0;
    int small_max_n = 1 << 16;  // largest request of the small class; 0: no small lane
    uint64_t next_ticket = 1;
    bool stop = false;
    int fail_batches = 0;
    // stats
    uint64_t submitted = 0, completed = 0, batches = 0, failed = 0;
    uint64_t batched_msms = 0;
    size_t peak_inflight = 0;
    int peak_busy = 0;
    std::vector<double> lat;
    std::thread th_sched;
    std::vector<std::thread> th_slot;  // one worker per slot
    std::vector<std::thread> th_comb;  // window combines
};

namespace {

size_t footprint_of(const gm::Plan& p, const std::vector<Request*>& rs) {
    size_t sc = 0;
    for (auto* r : rs) sc += (size_t)r->n * r->limbs * 8;
    return p.total + gm::al_size(sc);
}

bool plan_batch(gm_server* S, const std::vector<Request*>& rs, gm::Plan& p, std::vector<gm::MsmDesc>& desc,
                bool exact = true) {
    std::vector<gm::MsmSpec> sp(rs.size());
    int so = 0;
    for (size_t k = 0; k < rs.size(); ++k) {
        int c = 0, seg = 0;
        if (S->cfg.planner == GM_PLANNER_MLP) S->mlp.plan(rs[k]->curve, rs[k]->n, (int)rs.size(), rs[k]->bits, c, seg);
        sp[k] = {rs[k]->n, rs[k]->base_off, so, c, seg};
        so += rs[k]->n;
    }
    desc.resize(rs.size());
    return gm::make_plan(p, desc.data(), rs[0]->curve, rs[0]->limbs, rs[0]->bits, sp.data(), (int)rs.size(), exact);
}

// Slot buffers grow by a quarter beyond the request. Device buffers are freed
// and taken stream-ordered on the slot's stream, so a growing slot does not
// synchronise the device (and stall the other slots' batches).
hipError_t grow_dev(uint8_t*& ptr, size_t& cap, size_t want, hipStream_t s) {
    if (want <= cap) return hipSuccess;
    if (ptr) {
        GM_TRY_S(hipFreeAsync(ptr, s));
        ptr = nullptr;
        cap = 0;
    }
    want += want / 4;
    const hipError_t e = hipMallocAsync((void**)&ptr, want, s);
    if (e == hipSuccess)
        cap = want;
    else
        ptr = nullptr;
    return e;
}
hipError_t grow_pinned(uint8_t*& ptr, size_t& cap, size_t want) {
    if (want <= cap) return hipSuccess;
    if (ptr) {
        hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    want += want / 4;
    const hipError_t e = hipHostMalloc((void**)&ptr, want, hipHostMallocDefault);
    if (e == hipSuccess)
        cap = want;
    else
        ptr = nullptr;
    return e;
}

// The slot's pinned buffer: the error words, the window sums ERR_BYTES on,
// then the batch's descriptors.
size_t desc_at(const Slot& sl) { return gm::ERR_BYTES + gm::al_size(sl.res_bytes); }

template <class O>
hipError_t launch_slot(gm_server* S, Slot& sl) {
    const gm_bases* b = S->pool[sl.curve];
    std::vector<gmc::H2D> cp;
    cp.reserve(sl.reqs.size());
    uint8_t* sc = sl.d_sc;
    for (auto* r : sl.reqs) {
        const size_t bytes = (size_t)r->n * r->limbs * 8;
        cp.push_back({sc, r->borrowed ? r->borrowed : r->scalars.data(), bytes});
        sc += bytes;
    }
    GM_TRY_S(gmc::h2d_staged(sl.stage, (int)S->cfg.gpu_index, sl.stream, cp.data(), (int)cp.size()));
    GM_TRY_S(hipMemcpyAsync(sl.ws + sl.p.o_desc, sl.pinned + desc_at(sl), sl.desc.size() * sizeof(gm::MsmDesc),
                            hipMemcpyHostToDevice, sl.stream));
    GM_TRY_S(gm::Engine<O>::run(sl.stream, sl.p, reinterpret_cast<const Aff<O>*>(b->d),
                                reinterpret_cast<const uint64_t*>(sl.d_sc), sl.ws, nullptr));
    // error words and window sums in one copy (ERR_BYTES apart in the workspace)
    GM_TRY_S(hipMemcpyAsync(sl.pinned, sl.ws + sl.p.o_err, gm::ERR_BYTES + sl.res_bytes, hipMemcpyDeviceToHost,
                            sl.stream));
    return hipEventRecord(sl.ev, sl.stream);
}

bool is_large(const gm_server* S, const Request* r) { return S->small_max_n > 0 && r->n > S->small_max_n; }

// The request that heads the next batch and a free slot for it, or null:
// small requests first (FIFO among them); large ones while fewer than
// slots - 1 large batches are in flight (one slot stays with the small
// class; with one slot, or no small lane, every slot takes either).
const Request* pick_head(gm_server* S, int& sid) {
    sid = -1;
    int free_slot = -1, large_busy = 0;
    for (size_t i = 0; i < S->slots.size(); ++i) {
        if (!S->slots[i].busy) {
            if (free_slot < 0) free_slot = (int)i;
        } else if (S->slots[i].large) {
            ++large_busy;
        }
    }
    if (free_slot < 0) return nullptr;
    const Request* small = nullptr;
    const Request* big = nullptr;
    for (const Request* r : S->queue) {
        if (!is_large(S, r)) {
            small = r;
            break;
        }
        if (!big) big = r;
    }
    const int nslots = (int)S->slots.size();
    const int large_cap = (S->small_max_n > 0 && nslots > 1) ? nslots - 1 : nslots;
    const Request* head = small ? small : (big && large_busy < large_cap ? big : nullptr);
    if (head) sid = free_slot;
    return head;
}

void scheduler(gm_server* S) {
    hipSetDevice((int)S->cfg.gpu_index);
    for (;;) {
        std::unique_lock<std::mutex> lk(S->mu);
        int sid = -1;
        const Request* head = nullptr;
        S->cv_sched.wait(lk, [&] {
            if (S->stop) return true;
            head = pick_head(S, sid);
            return head != nullptr;
        });
        if (S->stop) return;
        // admission: FIFO over one (curve, limbs, bits, size) class, by
        // footprint (estimated plans during admission, the exact plan once at
        // the end)
        std::vector<Request*> take;
        gm::Plan p{};
        std::vector<gm::MsmDesc> desc;
        const bool large = is_large(S, head);
        const size_t budget = S->cfg.hbm_budget_bytes;
        std::vector<size_t> qidx;
        Request* reject = nullptr;
        size_t reject_qi = 0;
        for (size_t qi = 0; qi < S->queue.size() && (int)take.size() < S->cfg.max_batch; ++qi) {
            Request* r = S->queue[qi];
            if (r->curve != head->curve || r->limbs != head->limbs || r->bits != head->bits ||
                is_large(S, r) != large)
                continue;
            take.push_back(r);
            gm::Plan tp{};
            std::vector<gm::MsmDesc> td;
            const bool ok = plan_batch(S, take, tp, td, false);
            if (!ok) {
                take.pop_back();
                if (take.empty()) {
                    reject = r;  // cannot be planned at all
                    reject_qi = qi;
                }
                break;
            }
            if (S->inflight_bytes + footprint_of(tp, take) > budget) {
                take.pop_back();
                if (take.empty() && S->inflight_bytes == 0) {
                    take.push_back(r);  // larger than the budget: runs alone on an idle device
                    qidx.push_back(qi);
                }
                break;
            }
            qidx.push_back(qi);
        }
        if (reject) {
            S->queue.erase(S->queue.begin() + (long)reject_qi);
            reject->rc = -1;
            reject->done = true;
            reject->t_done = now_ms();
            S->completed += 1;
            S->failed += 1;
            S->cv_done.notify_all();
            continue;
        }
        if (take.empty()) {
            // wait for an in-flight batch to release HBM
            S->cv_sched.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        if (!plan_batch(S, take, p, desc, true)) {
            S->cv_sched.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        for (auto it = qidx.rbegin(); it != qidx.rend(); ++it) S->queue.erase(S->queue.begin() + (long)*it);
        Slot& sl = S->slots[sid];
        sl.busy = true;
        sl.launched = false;
        sl.large = large;
        sl.reqs = take;
        sl.p = p;
        sl.desc = desc;
        sl.curve = head->curve;
        sl.footprint = footprint_of(p, take);
        sl.fail = S->fail_batches > 0;
        if (sl.fail) --S->fail_batches;
        S->inflight_bytes += sl.footprint;
        S->peak_inflight = std::max(S->peak_inflight, S->inflight_bytes);
        int busy = 0;
        for (auto& s2 : S->slots) busy += s2.busy;
        S->peak_busy = std::max(S->peak_busy, busy);
        // the slot's worker stages, launches and completes the batch
        S->cv_slot.notify_all();
    }
}

// Publishes one finished request (caller holds S->mu).
void publish(gm_server* S, Request* r, int rc, double t) {
    r->rc = rc;
    r->done = true;
    r->t_done = t;
    r->scalars.clear();
    r->scalars.shrink_to_fit();
    S->completed += 1;
    if (rc == 0)
        S->lat.push_back(t - r->t_submit);
    else
        S->failed += 1;
}

// Device side of one batch on its slot: buffers, staged scalars, pipeline,
// window sums back to the slot's pinned buffer, the slot's event.
hipError_t launch_batch(gm_server* S, Slot& sl) {
    if (sl.fail) return hipEventRecord(sl.ev, sl.stream);  // injected failure: nothing is launched
    size_t sc_bytes = 0;
    for (auto* r : sl.reqs) sc_bytes += (size_t)r->n * r->limbs * 8;
    GM_TRY_S(grow_dev(sl.ws, sl.ws_cap, sl.p.total, sl.stream));
    GM_TRY_S(grow_dev(sl.d_sc, sl.sc_cap, sc_bytes, sl.stream));
    GM_TRY_S(grow_pinned(sl.pinned, sl.pinned_cap, desc_at(sl) + sl.desc.size() * sizeof(gm::MsmDesc)));
    memcpy(sl.pinned + desc_at(sl), sl.desc.data(), sl.desc.size() * sizeof(gm::MsmDesc));
    return sl.curve == GM_CURVE_G1 ? launch_slot<FpOps>(S, sl) : launch_slot<Fp2Ops>(S, sl);
}

// Worker of one slot: launches each batch the scheduler assigns to it, waits
// for the device, releases the slot and hands the window combines to the
// combine threads, in whatever order the slots finish.
void slot_worker(gm_server* S, int sid) {
    hipSetDevice((int)S->cfg.gpu_index);
    Slot& sl = S->slots[sid];
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_slot.wait(lk, [&] { return S->stop || (sl.busy && !sl.launched); });
            if (!(sl.busy && !sl.launched)) return;  // stop with nothing assigned to this slot
        }
        sl.res_bytes = (size_t)sl.p.win_total * gm::xyzz_bytes(sl.curve);
        if (launch_batch(S, sl) != hipSuccess) {
            // Copies and kernels of this batch may already be queued on the
            // slot's stream: the slot (and its HBM share) is released only
            // after they have drained, never on the event of the slot's
            // previous batch.
            sl.fail = true;
            (void)hipGetLastError();
            if (hipEventRecord(sl.ev, sl.stream) != hipSuccess) (void)hipStreamSynchronize(sl.stream);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            sl.launched = true;
            S->batches += 1;
            S->batched_msms += sl.reqs.size();
        }
        const hipError_t e = hipEventSynchronize(sl.ev);
        int rc = 0;
        if (sl.fail || e != hipSuccess) {
            rc = -4;
        } else {
            uint32_t err;
            memcpy(&err, sl.pinned, 4);
            if (err) rc = -5;
        }
        std::vector<CombineJob> jobs;
        if (rc == 0) {
            const uint32_t* raw = reinterpret_cast<const uint32_t*>(sl.pinned + gm::ERR_BYTES);
            const size_t words = gm::xyzz_bytes(sl.curve) / 4;
            jobs.resize(sl.reqs.size());
            for (size_t j = 0; j < sl.reqs.size(); ++j) {
                const gm::MsmDesc& d = sl.desc[j];
                CombineJob& job = jobs[j];
                job.r = sl.reqs[j];
                job.curve = sl.curve;
                job.nw = (int)d.nw;
                job.c = (int)d.c;
                job.raw.assign(raw + (size_t)d.win_base * words, raw + (size_t)(d.win_base + d.nw) * words);
            }
        }
        const double t = now_ms();
        std::lock_guard<std::mutex> lk(S->mu);
        if (rc != 0)
            for (auto* r : sl.reqs) publish(S, r, rc, t);
        // the small class's combines go first, ahead of queued large ones
        if (sl.large) {
            for (auto& job : jobs) S->jobs.push_back(std::move(job));
        } else {
            for (auto it = jobs.rbegin(); it != jobs.rend(); ++it) S->jobs.push_front(std::move(*it));
        }
        sl.reqs.clear();
        sl.busy = false;
        sl.launched = false;
        S->inflight_bytes -= sl.footprint;
        S->cv_job.notify_all();
        S->cv_done.notify_all();
        S->cv_sched.notify_one();
    }
}

// Combine thread: Horner window combine of one request at a time.
void combiner(gm_server* S) {
    for (;;) {
        CombineJob job;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_job.wait(lk, [&] { return S->stop || !S->jobs.empty(); });
            if (S->jobs.empty()) return;  // stop with nothing left to combine
            job = std::move(S->jobs.front());
            S->jobs.pop_front();
            S->combining += 1;
        }
        if (job.curve == GM_CURVE_G1) {
            gmh::Fp x, y;
            const bool f = gmc::host_combine<gmh::Fp>(x, y, job.raw.data(), job.nw, job.c);
            gmc::write_affine_abi<gmh::Fp>(job.r->result, f, x, y);
        } else {
            gmh::Fp2 x, y;
            const bool f = gmc::host_combine<gmh::Fp2>(x, y, job.raw.data(), job.nw, job.c);
            gmc::write_affine_abi<gmh::Fp2>(job.r->result, f, x, y);
        }
        const double t = now_ms();
        std::lock_guard<std::mutex> lk(S->mu);
        publish(S, job.r, 0, t);
        S->combining -= 1;
        S->cv_done.notify_all();
    }
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int gm_server_create(gm_server** out, const gm_server_config* cfg, const gm_bases* g1_pool, const gm_bases* g2_pool) {
    if (!out || !cfg || (!g1_pool && !g2_pool)) return -1;
    if (cfg->slots < 1 || cfg->slots > 16 || cfg->max_batch < 1) return -1;
    if ((g1_pool && g1_pool->curve != GM_CURVE_G1) || (g2_pool && g2_pool->curve != GM_CURVE_G2)) return -1;
    if (hipSetDevice((int)cfg->gpu_index) != hipSuccess) return -2;
    gm_server* S = new gm_server();
    S->cfg = *cfg;
    S->pool[GM_CURVE_G1] = g1_pool;
    S->pool[GM_CURVE_G2] = g2_pool;
    if (cfg->planner == GM_PLANNER_MLP) {
        if (!cfg->mlp_params) {
            delete S;
            return -1;
        }
        S->mlp.P.assign(cfg->mlp_params, cfg->mlp_params + GM_MLP_PARAMS);
    }
    S->slots.resize((size_t)cfg->slots);
    for (auto& sl : S->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            gm_server_destroy(S);
            return -2;
        }
    }
    if (cfg->small_max_n > 0)
        S->small_max_n = cfg->small_max_n;
    else if (cfg->small_max_n < 0)
        S->small_max_n = 0;
    S->th_sched = std::thread(scheduler, S);
    for (int i = 0; i < cfg->slots; ++i) S->th_slot.emplace_back(slot_worker, S, i);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    for (unsigned i = 0; i < std::min(8u, hw); ++i) S->th_comb.emplace_back(combiner, S);
    *out = S;
    return 0;
}

int gm_server_submit(gm_server* S, int curve, int base_offset, int n, const uint64_t* scalars, int limbs,
                     int scalar_bits, uint32_t flags, uint64_t* ticket) {
    const double t0 = now_ms();
    if (!S || !scalars || !ticket || n <= 0 || limbs < 1 || limbs > 8 || base_offset < 0) return -1;
    if (curve != GM_CURVE_G1 && curve != GM_CURVE_G2) return -1;
    const gm_bases* b = S->pool[curve];
    if (!b || (long)base_offset + n > b->n) return -1;
    Request* r = new Request();
    r->curve = curve;
    r->base_off = base_offset;
    r->n = n;
    r->limbs = limbs;
    r->bits = scalar_bits;
    if (flags & GM_SUBMIT_BORROW)
        r->borrowed = scalars;
    else
        r->scalars.assign(scalars, scalars + (size_t)n * limbs);
    r->t_submit = t0;
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
        delete r;
        return -1;
    }
    r->ticket = S->next_ticket++;
    S->table[r->ticket] = r;
    S->queue.push_back(r);
    S->submitted += 1;
    *ticket = r->ticket;
    S->cv_sched.notify_one();
    return 0;
}

int gm_server_wait(gm_server* S, uint64_t ticket, void* result_affine, double* latency_ms, int timeout_ms) {
    if (!S || !result_affine) return -1;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->table.find(ticket);
    if (it == S->table.end()) return -1;
    Request* r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, pred);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return -6;
    const int rc = r->rc;
    if (rc == 0) memcpy(result_affine, r->result, r->curve == GM_CURVE_G1 ? 120 : 232);
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    S->table.erase(it);
    delete r;
    return rc;
}

int gm_server_get_stats(gm_server* S, gm_server_stats* out) {
    if (!S || !out) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    memset(out, 0, sizeof *out);
    out->submitted = S->submitted;
    out->completed = S->completed;
    out->failed = S->failed;
    out->batches = S->batches;
    out->mean_batch = S->batches ? (double)S->batched_msms / (double)S->batches : 0.0;
    out->p50_ms = percentile(S->lat, 0.50);
    out->p99_ms = percentile(S->lat, 0.99);
    out->max_ms = S->lat.empty() ? 0.0 : *std::max_element(S->lat.begin(), S->lat.end());
    out->peak_inflight_bytes = S->peak_inflight;
    out->peak_concurrent_batches = S->peak_busy;
    return 0;
}

int gm_server_inject_failures(gm_server* S, int batches) {
    if (!S || batches < 0) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    S->fail_batches = batches;
    return 0;
}

void gm_server_destroy(gm_server* S) {
    if (!S) return;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        // drain: wait until queued, in-flight and combining work has completed
        S->cv_done.wait(lk, [&] {
            if (!S->queue.empty() || !S->jobs.empty() || S->combining > 0) return false;
            for (auto& sl : S->slots)
                if (sl.busy) return false;
            return true;
        });
        S->stop = true;
        S->cv_sched.notify_all();
        S->cv_slot.notify_all();
        S->cv_job.notify_all();
    }
    if (S->th_sched.joinable()) S->th_sched.join();
    for (auto& th : S->th_slot)
        if (th.joinable()) th.join();
    for (auto& th : S->th_comb)
        if (th.joinable()) th.join();
    hipSetDevice((int)S->cfg.gpu_index);
    for (auto& sl : S->slots) {
        if (sl.stream) {
            if (sl.ws) hipFreeAsync(sl.ws, sl.stream);
            if (sl.d_sc) hipFreeAsync(sl.d_sc, sl.stream);
            hipStreamSynchronize(sl.stream);
        }
        if (sl.ev) hipEventDestroy(sl.ev);
        if (sl.stream) hipStreamDestroy(sl.stream);
        if (sl.pinned) hipHostFree(sl.pinned);
        gmc::staging_release(sl.stage);
    }
    for (auto& kv : S->table) delete kv.second;
    delete S;
}

}  // extern "C"
