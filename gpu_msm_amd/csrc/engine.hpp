// Host-side interface of the MSM engine: plans, workspace layout and the
// per-curve kernel drivers (instantiated in kern_g1.hip / kern_g2.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "curve.hpp"

namespace gm {

enum Curve { CURVE_G1 = 1, CURVE_G2 = 2 };

// Most u64 words per scalar the digit kernels accept.
constexpr int MAX_SCALAR_LIMBS = 8;

// Reference ABI record sizes (reference include/curve.h:13-43).
template <class O>
struct AbiLayout;
template <>
struct AbiLayout<FpOps> {
    static constexpr int AFF_BYTES = 120;   // G1Point
    static constexpr int PROJ_BYTES = 168;  // G1ProjectivePoint
    static constexpr int ELEM_U64 = 7;
};
template <>
struct AbiLayout<Fp2Ops> {
    static constexpr int AFF_BYTES = 232;   // G2Point
    static constexpr int PROJ_BYTES = 336;  // G2ProjectivePoint
    static constexpr int ELEM_U64 = 14;
};

size_t al_size(size_t x);

// Profiling stages recorded with HIP events on the MSM's own stream.
enum Stage { ST_START = 0, ST_HIST, ST_SCAN, ST_SCATTER, ST_ACCUM, ST_FIXUP, ST_REDUCE, ST_WSUM, ST_COUNT };

struct Prof {
    hipEvent_t ev[ST_COUNT];
    bool on;
};

// One MSM of a batch, as the kernels see it (80 B, device resident).
struct MsmDesc {
    uint32_t n;          // points
    uint32_t base_off;   // first base in the resident pool
    uint32_t scal_off;   // first scalar (in scalars, each `L` limbs)
    uint32_t c;          // window bits
    uint32_t nw;         // windows
    uint32_t B;          // buckets per window = 2^(c-1)
    uint32_t seg;        // buckets per bucket-reduce lane
    uint32_t nseg;       // segments per window = ceil(B / seg); the last may be shorter
    uint32_t bkt_base;   // first global bucket
    uint32_t pt_base;    // first flattened point
    uint32_t rank_base;  // first (window, point) rank slot
    uint32_t seg_base;   // first global segment
    uint32_t win_base;   // first global window
    uint32_t chunk;      // entries per accumulation lane for this MSM
    uint32_t lane_base;  // first accumulation lane
    uint32_t tile_base;  // first window-sum tile (WSUM_TILE segments each)
    uint32_t stile_base; // first sort tile (stile_pts points of one window)
    uint32_t stile_pts;  // points per sort tile
    uint32_t tb_base;    // first per-tile bucket base (B words per sort tile)
    uint32_t flags;      // DESC_WSUM_DIRECT: window sums by one lane per window
};
static_assert(sizeof(MsmDesc) == 80, "MsmDesc layout");

// Bytes from the error words (error flag, heavy-bucket count, entry count)
// to the window sums in a plan's workspace: one device-to-host copy of
// ERR_BYTES + the window sums brings the flags and the sums together.
constexpr size_t ERR_BYTES = 256;

// Segments summed by one window-sum tile (64 threads x 4) before the
// per-window total over tiles.
constexpr uint32_t WSUM_TILE = 256;
// Windows of at most 4 segments, or of at most WSUM_DIRECT_MAX in batches of
// many windows, are summed by one lane each (k_window_direct): a sequential
// sum of a few segments is no deeper than the two tree stages, and a batch
// of many small MSMs then runs one lane, not one workgroup, per window.
constexpr uint32_t WSUM_DIRECT_MAX = 24;
constexpr uint32_t DESC_WSUM_DIRECT = 1u;

// Counting sort with per-tile bucket counters in LDS: used while every MSM of
// a batch has c <= SORT_LDS_MAX_C (2^(c-1) counters of 4 bytes per
// workgroup); larger windows fall back to one global atomic per entry.
constexpr int SORT_LDS_MAX_C = 16;
// Points of one window counted by one sort workgroup.
constexpr uint32_t SORT_TILE_PTS = 1u << 17;
// Buckets per block of the two-pass placement (windows of more buckets than
// this place by blocks, then by bucket).
constexpr uint32_t SORT_BLOCK = 128;

// What a caller asks for: c <= 0 and seg <= 0 pick the defaults.
struct MsmSpec {
    int n;
    int base_off;
    int scal_off;
    int c;
    int seg;
};

struct Plan {
    int curve;
    int L;          // u64 limbs per scalar (all MSMs of a batch)
    int bits;       // scalar bit bound (values must be < 2^bits)
    int count;      // MSMs in the batch
    uint32_t n_total, ent_total, bkt_total, seg_total, win_total, tile_total;
    uint32_t stile_total;  // sort tiles (0: the global-atomic fallback sort)
    uint32_t lds_B;        // largest bucket count of the batch (LDS counters per sort tile)
    uint32_t qmax;         // two-pass placement: block slots per sort tile (0: one pass)
    int lanes;      // accumulation lanes (upper bound: all digits non-zero)
    bool tail_split; // bucket reduction and window sums on split lane groups (FpPairOps / Fp2QuadOps)
    bool fixup_split; // the fix-up's serial sums on split lane groups
    size_t cub_bytes;
    // workspace offsets (bytes)
    size_t o_desc, o_cnt, o_off, o_rank, o_dig, o_tbase, o_blk, o_qbase, o_tmp, o_sorted, o_buckets, o_head, o_tail,
        o_heavy, o_segs, o_tiles, o_wsum, o_out, o_err,
        o_cub,
        total;
    MsmDesc* host_desc;  // owned by the caller of make_plan (count entries)
};

size_t xyzz_bytes(int curve);
size_t aff_bytes(int curve);
int choose_window(int curve, int n, int bits);
bool window_ok(int bits, int c);
int choose_seg(int B);
// Fills p (and p.host_desc[0..count)) for a batch; false on invalid specs.
// exact_scan_bytes = false bounds the scan's temporary storage instead of
// asking the device (fast, for admission estimates).
bool make_plan(Plan& p, MsmDesc* desc_storage, int curve, int L, int bits, const MsmSpec* specs, int count,
               bool exact_scan_bytes = true);

template <class O>
struct Engine {
    // Runs digits -> sort -> accumulate -> reduce -> window sums into the
    // workspace (window sums at p.o_wsum, MSM j's at win_base_j). The
    // descriptors must already be at p.o_desc. Asynchronous on `s`.
    static hipError_t run(hipStream_t s, const Plan& p, const Aff<O>* d_bases, const uint64_t* d_scalars, uint8_t* d_ws,
                          Prof* prof);
    // Horner over window sums on the device, one lane per MSM; canonical
    // affine results at p.o_out (count entries).
    static hipError_t device_combine(hipStream_t s, const Plan& p, uint8_t* d_ws);
    static hipError_t import_aff(hipStream_t s, Aff<O>* out, const void* d_abi, int n, bool to_mont);
    static hipError_t export_aff(hipStream_t s, void* d_abi, const Aff<O>* in, int n);
    static hipError_t gen_bases(hipStream_t s, Aff<O>* out, const uint64_t* d_k, int n, const Aff<O>& g);
    static hipError_t aff_to_abi_proj(hipStream_t s, void* d_abi_proj, const Aff<O>* in);
};

extern template struct Engine<FpOps>;
extern template struct Engine<Fp2Ops>;

// Per-stage launchers, each compiled in its own translation unit so the
// heavy G2 instantiations build in parallel.
hipError_t launch_sort(hipStream_t s, const Plan& p, const uint64_t* d_scalars, uint8_t* ws, Prof* prof);
template <class O>
hipError_t launch_accumulate(hipStream_t s, const Plan& p, const Aff<O>* d_bases, uint8_t* ws);
template <class O>
hipError_t launch_fixup(hipStream_t s, const Plan& p, uint8_t* ws);
template <class O>
hipError_t launch_reduce(hipStream_t s, const Plan& p, uint8_t* ws);
template <class O>
hipError_t launch_window_sum(hipStream_t s, const Plan& p, uint8_t* ws);
template <class O>
hipError_t launch_horner(hipStream_t s, const MsmDesc* desc, int count, const Xyzz<O>* wsum, Aff<O>* out);
template <class O>
hipError_t launch_import(hipStream_t s, Aff<O>* out, const void* d_abi, int n, bool to_mont);
template <class O>
hipError_t launch_export(hipStream_t s, void* d_abi, const Aff<O>* in, int n);
template <class O>
hipError_t launch_gen_bases(hipStream_t s, Aff<O>* out, const uint64_t* d_k, int n, const Aff<O>& g);
template <class O>
hipError_t launch_aff_to_abi_proj(hipStream_t s, void* d_abi_proj, const Aff<O>* in, int count);
hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);

}  // namespace gm
