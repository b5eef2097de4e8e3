// Batch plans: per-MSM window choice, descriptor prefix sums and the
// workspace layout of one pipeline run.
#include <hipcub/hipcub.hpp>

#include <string.h>

#include <algorithm>

#include "engine.hpp"

namespace gm {

size_t xyzz_bytes(int curve) { return curve == CURVE_G1 ? sizeof(Xyzz<FpOps>) : sizeof(Xyzz<Fp2Ops>); }
size_t aff_bytes(int curve) { return curve == CURVE_G1 ? sizeof(Aff<FpOps>) : sizeof(Aff<Fp2Ops>); }
size_t al_size(size_t x) { return (x + 255) & ~(size_t)255; }

static int windows_for(int bits, int c) { return (bits + 1 + c - 1) / c; }

// A width whose top window keeps fewer than 4/5 of its c bits sends every
// point of that window to a handful of buckets (c = 11 for 299-bit scalars
// leaves it 3 bits: 2^14 points land in 4 buckets). The bucket phase then
// sums hundreds of chunk partials per bucket on a long serial chain, which
// the sweeps measure at up to 1.6x the time of a neighbouring width.
bool window_ok(int bits, int c) {
    const int top = bits + 1 - (windows_for(bits, c) - 1) * c;
    return 5 * top >= 4 * c;
}

// Buckets per bucket-reduce lane assumed by the window cost model (the plan
// itself sizes segments for the whole batch, see make_plan).
int choose_seg(int B) { return B < 8 ? B : 8; }

// Relative cost in Fp-multiplication equivalents: mixed adds of the bucket
// phase (10 products each, 3x for Fp2), the running-sum adds (two full adds
// per bucket, weighted up for their lower occupancy) and per-segment scalar
// multiples of the bucket reduction, and a per-(window, point) charge for
// recoding and sorting.
static double plan_cost(int curve, int n, int bits, int c) {
    const int nw = windows_for(bits, c);
    const double B = (double)(1 << (c - 1));
    const double seg = choose_seg(1 << (c - 1));
    const double fm = curve == CURVE_G1 ? 1.0 : 3.0;
    const double accum = (double)nw * n * 10.5 * fm;
    const double reduce = (double)nw * (B * 56.0 + (B / seg) * 18.0 * c) * fm;
    const double sort = (double)nw * n * 0.6;
    return accum + reduce + sort;
}

int choose_window(int curve, int n, int bits) {
    int best = 2;
    double bc = 1e300;
    for (int c = 2; c <= 20; ++c) {
        if (!window_ok(bits, c)) continue;
        const double v = plan_cost(curve, n, bits, c);
        if (v < bc) {
            bc = v;
            best = c;
        }
    }
    return best;
}

bool make_plan(Plan& p, MsmDesc* d, int curve, int L, int bits, const MsmSpec* specs, int count,
               bool exact_scan_bytes) {
    memset(&p, 0, sizeof p);
    if (count < 1 || L < 1 || L > 8 || !d || !specs) return false;
    if (bits <= 0 || bits > 64 * L) bits = 64 * L;
    p.curve = curve;
    p.L = L;
    p.bits = bits;
    p.count = count;
    p.host_desc = d;
    // Buckets per bucket-reduction lane for the MSMs that leave it to the
    // plan: the smallest count >= 4 that keeps the batch's reduction within
    // REDUCE_THREADS threads, i.e. about one wave per SIMD (the reduction
    // runs at one wave per SIMD, so a second round would double its time,
    // while fewer waves leave SIMDs idle and lengthen each lane's running
    // sums), capped where a batch needs more than one round (16 for G1, 32
    // for G2). A segment takes one thread (G1) or a lane pair (G2), doubled
    // when the tail kernels split each point operation's products over two
    // such groups (tail_split): the plan splits while the batch's segments
    // then still fit one round, and not past it, where the reduction is
    // bound by throughput and the unsplit kernels do less work. Small MSMs
    // get 4, the sweeps' optimum for them; at N = 2^20 G1 gets 11 and G2 21.
    constexpr uint64_t REDUCE_THREADS = 62 * 1024;
    uint64_t auto_bkt = 0, fixed_segs = 0;
    for (int j = 0; j < count; ++j) {
        const MsmSpec& s = specs[j];
        const int c = s.c > 0 ? s.c : choose_window(curve, s.n, bits);
        if (c < 2 || c > 24) return false;
        const uint64_t B = 1ull << (c - 1), nbk = B * (uint64_t)windows_for(bits, c);
        if (s.seg > 0) {
            const uint64_t sg = std::min<uint64_t>(B, (uint64_t)s.seg);
            fixed_segs += (B + sg - 1) / sg * windows_for(bits, c);
        } else {
            auto_bkt += nbk;
        }
    }
    const uint64_t seg_cap = curve == CURVE_G2 ? 32 : 16;
    auto seg_for = [&](uint64_t lpi) {
        uint64_t sg = 4;
        if (fixed_segs * lpi < REDUCE_THREADS)
            sg = std::max<uint64_t>(sg, (auto_bkt * lpi + (REDUCE_THREADS - fixed_segs * lpi) - 1) /
                                            (REDUCE_THREADS - fixed_segs * lpi));
        return sg;
    };
    const uint64_t group = curve == CURVE_G2 ? 2 : 1;  // threads per segment, unsplit
    uint64_t auto_seg = seg_for(2 * group);
    p.tail_split = true;
    if (auto_seg > seg_cap) {
        p.tail_split = false;
        auto_seg = seg_for(group);
    }
    auto_seg = std::min(auto_seg, seg_cap);
    uint64_t pt = 0, rank = 0, bkt = 0, seg = 0, win = 0, tile = 0;
    for (int j = 0; j < count; ++j) {
        const MsmSpec& s = specs[j];
        if (s.n < 0 || s.base_off < 0 || s.scal_off < 0) return false;
        const int c = s.c > 0 ? s.c : choose_window(curve, s.n, bits);
        if (c < 2 || c > 24) return false;
        const int nw = windows_for(bits, c);
        const int B = 1 << (c - 1);
        int sg = s.seg > 0 ? s.seg : (int)auto_seg;
        if (sg > B) sg = B;
        MsmDesc& m = d[j];
        memset(&m, 0, sizeof m);
        m.n = (uint32_t)s.n;
        m.base_off = (uint32_t)s.base_off;
        m.scal_off = (uint32_t)s.scal_off;
        m.c = (uint32_t)c;
        m.nw = (uint32_t)nw;
        m.B = (uint32_t)B;
        m.seg = (uint32_t)sg;
        m.nseg = (uint32_t)((B + sg - 1) / sg);  // the last segment may be shorter
        m.bkt_base = (uint32_t)bkt;
        m.pt_base = (uint32_t)pt;
        m.rank_base = (uint32_t)rank;
        m.seg_base = (uint32_t)seg;
        m.win_base = (uint32_t)win;
        pt += (uint64_t)s.n;
        rank += (uint64_t)nw * s.n;
        bkt += (uint64_t)nw * B;
        seg += (uint64_t)nw * m.nseg;
        win += (uint64_t)nw;
        if (rank >= 0xFFFFFFFFull || bkt >= 0x7FFFFFFFull) return false;
    }
    // Window sums: one lane per window (k_window_direct) for windows of at
    // most 4 segments, and up to WSUM_DIRECT_MAX when the batch has so many
    // windows that one workgroup per window would need more than one round
    // of the chip's SIMDs; otherwise tiles and a tree per window.
    for (int j = 0; j < count; ++j) {
        MsmDesc& m = d[j];
        const bool direct = m.nseg <= 4 || (m.nseg <= WSUM_DIRECT_MAX && win >= 1024);
        m.flags = direct ? DESC_WSUM_DIRECT : 0u;
        m.tile_base = (uint32_t)tile;
        if (!direct) tile += (uint64_t)m.nw * ((m.nseg + WSUM_TILE - 1) / WSUM_TILE);
    }
    // Sort tiles: every MSM of the batch must fit its counters in LDS.
    bool lds_sort = true;
    uint32_t maxB = 0;
    for (int j = 0; j < count; ++j) {
        lds_sort = lds_sort && (int)d[j].c <= SORT_LDS_MAX_C;
        maxB = std::max(maxB, d[j].B);
    }
    uint64_t stile = 0, tb = 0;
    if (lds_sort) {
        for (int j = 0; j < count; ++j) {
            MsmDesc& m = d[j];
            m.stile_pts = std::max<uint32_t>(1, std::min<uint32_t>(m.n, SORT_TILE_PTS));
            const uint64_t per_w = m.n ? ((uint64_t)m.n + m.stile_pts - 1) / m.stile_pts : 0;
            m.stile_base = (uint32_t)stile;
            m.tb_base = (uint32_t)tb;
            stile += per_w * m.nw;
            tb += per_w * m.nw * m.B;
        }
        if (tb >= 0xFFFFFFFFull) return false;
    }
    // split tail kernels only while all segments' lane groups fit one round
    // of one wave per SIMD (1024 SIMDs x 64 lanes)
    if (p.tail_split && 2 * group * seg > 64 * 1024) p.tail_split = false;
    // The fix-up sums each cut bucket's chunk partials serially, one lane (G1)
    // or lane pair (G2) per bucket. Few buckets leave it below one wave per
    // SIMD, where one XYZZ addition costs its full latency (~65 us for G1),
    // so it splits each addition's products over two lane groups while the
    // split launch fits two rounds of two waves per SIMD (256K threads):
    // G1 2^12-2^18 3-7% and G2 2^12-2^16 1.5-7% faster per MSM; past it
    // (G1 and G2 2^19, 2^20) the split costs 1% or nothing.
    p.fixup_split = bkt * 2 * group <= 256 * 1024;
    p.stile_total = (uint32_t)stile;
    p.lds_B = lds_sort ? maxB : 0;
    // Windows of more than SORT_BLOCK buckets place their entries in two
    // passes (by blocks of SORT_BLOCK buckets, then by bucket), each staged in
    // LDS so that whole runs are stored: one pass with a counter per bucket
    // measured 0.30 ms at G1 2^20, of which 0.25 ms were its scattered stores.
    p.qmax = (lds_sort && maxB > SORT_BLOCK) ? maxB / SORT_BLOCK + 2 : 0;
    p.n_total = (uint32_t)pt;
    p.ent_total = (uint32_t)rank;
    p.bkt_total = (uint32_t)bkt;
    p.seg_total = (uint32_t)seg;
    p.win_total = (uint32_t)win;
    p.tile_total = (uint32_t)tile;
    // Entries per accumulation lane, per MSM (chunks never straddle two
    // MSMs). Parallelism asks for at least ~2^17 lanes over the whole batch
    // (one round of two waves per SIMD for G1, two rounds of lane pairs for
    // G2): chunk 40 for big G1 runs, 80 for G2, whose fix-up adds cost three
    // times more and whose longer mixed adds hide the gathers with fewer
    // lanes, halved down to 2 for smaller runs, so the serial depth of one
    // mixed add per entry stays short; the fix-up asks that a typical bucket
    // of the MSM spans at most ~4 chunks, so chunk >= the MSM's average
    // bucket / 4. At N = 2^20 chunks of 40 (G1) and 80 (G2) make exactly four
    // rounds of two waves per SIMD and cut a fifth fewer buckets than 32 / 64:
    // accumulation + fix-up 4.00-4.03 -> 3.94 ms (G1), 13.58-13.79 ->
    // 13.43 ms (G2); longer chunks cut fewer buckets still but slow the
    // accumulation more than that (80-160 for G1 and 160 for G2 measured the
    // same or slower). Below 2^19 points the lane count decides: asking for
    // 2^16 lanes instead (half as many) measured G2 2^18 6.78-6.83 ms against
    // 5.66-5.70, G2 2^17 3.93 against 3.46 and G1 2^17 1.68 against 1.56.
    int k_par = curve == CURVE_G2 ? 80 : 40;
    while (k_par > 2 && rank < (uint64_t)k_par * 131072) k_par >>= 1;
    uint64_t lanes = 0;
    for (int j = 0; j < count; ++j) {
        MsmDesc& m = d[j];
        const uint64_t avg = (uint64_t)m.n / m.B;
        uint32_t k_span = 2;
        while (k_span < 1024 && (uint64_t)k_span * 4 < avg) k_span <<= 1;
        m.chunk = std::max<uint32_t>((uint32_t)k_par, k_span);
        m.lane_base = (uint32_t)lanes;
        lanes += ((uint64_t)m.nw * m.n + m.chunk - 1) / m.chunk;
    }
    if (lanes >= 0x7FFFFFFFull) return false;
    p.lanes = (int)lanes;
    size_t cub = 0;
    if (!exact_scan_bytes ||
        hipcub::DeviceScan::ExclusiveSum(nullptr, cub, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)(bkt + 1)) !=
            hipSuccess) {
        // No device to ask (plan queries on a host without a GPU): a bound on
        // the decoupled-lookback scan state.
        cub = (bkt / 128 + 1024) * 16 + 65536;
    }
    p.cub_bytes = cub;
    const size_t xb = xyzz_bytes(curve);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += al_size(bytes);
        return at;
    };
    p.o_desc = take((size_t)count * sizeof(MsmDesc));
    p.o_cnt = take((bkt + 1) * 4);
    p.o_off = take((bkt + 1) * 4);
    // entry ranks (global-atomic sort) or digit codes + per-tile bucket bases
    p.o_rank = take(p.lds_B ? 0 : rank * 4 + 4);
    p.o_dig = take(p.lds_B ? rank * 2 + 4 : 0);
    p.o_tbase = take(p.lds_B && !p.qmax ? tb * 4 + 4 : 0);
    p.o_blk = take(p.qmax ? ((bkt + SORT_BLOCK - 1) / SORT_BLOCK + 1) * 4 : 0);
    p.o_qbase = take(p.qmax ? stile * p.qmax * 4 + 4 : 0);
    p.o_tmp = take(p.qmax ? rank * 8 + 8 : 0);
    p.o_sorted = take(rank * 8 + 8);
    p.o_buckets = take(bkt * xb);
    p.o_head = take((size_t)p.lanes * xb);
    p.o_tail = take((size_t)p.lanes * xb);
    p.o_heavy = take((bkt + 1) * 4);
    p.o_segs = take(seg * xb);
    p.o_tiles = take(tile * xb);
    // error flag, heavy-bucket count and entry count, right before the window
    // sums (ERR_BYTES apart) so that one copy brings both to the host
    p.o_err = take(16);
    p.o_wsum = take(win * xb);
    p.o_out = take((size_t)count * aff_bytes(curve));
    if (p.o_wsum != p.o_err + ERR_BYTES) return false;
    p.o_cub = take(cub);
    p.total = o;
    return true;
}

}  // namespace gm
