// device.h -- device-side layout constants, the panel object and the
// per-call workspace shared by the engine and the request server.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "internal.h"

namespace ldsc {

// Tiled operand layout (built per call from the resident .bed slabs):
//  * SNPs in groups of 32 (one MFMA row/column block); band tiles are 64x64.
//  * Individuals in K-blocks of 128.  Tile (group g, K-block kb) is 1 KiB:
//    64 lanes x 16 bytes.  Lane l holds SNP 32g + l%32; its dword t (0..3)
//    holds the 16 individuals kb*128 + t*32 + (l/32)*16 + e, e = 4f + k,
//    with the 2-bit PLINK code of individual e at bits 8k + 2f.  A wave
//    loads a whole tile with one coalesced 16-byte load per lane.
//  * Padding individuals and SNPs carry code 01 (missing): they add nothing
//    to any integer product and are masked out of every pair.
constexpr int kTile = 64;
constexpr int kGroup = 32;
constexpr int kKBlock = 128;
// Largest N of the e2m1 exact kernel: its f32 accumulators hold sums of
// products of v/2 and q/2 (v <= 3), multiples of 1/4 up to 2.25 N_pad,
// exactly while 9 N_pad <= 2^24.
constexpr uint64_t kFp4MaxIndiv = 1800000;

inline uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Arguments of one call's band contraction (kernel 3, band_common.h).
struct BandArgs {
    const u32x4* tiled;
    const int2* tiles;
    const uint32_t* lo;
    const int2* cs;
    const double* rs;
    const float4* lut;
    const double* aeff;
    double* rowp;
    double* colp;
    const float* sk;       // sketched columns (CountSketch path) or nullptr
    const float* sk_scale; // their per-SNP renorm scales
    const double* kappa;   // per-SNP kurtosis of the MAF-aware sketch or nullptr
    // integer-operand sketch form: W and E int8 tiles, and per SNP (m_pad
    // each, in this order) the scale inv*sc, delta and A (sketch.hip)
    const u32x4* skw;
    const u32x4* ske;
    const double* sk8;
    double sk_tt;
    uint32_t nkb, ntiles, per_xcd, m, m_pad, K;
    uint32_t tile0;        // first tile of this launch in the call's tile list
    double r2u_a, r2u_b, n_inv_sq_r2u_a, n_inv_sq, d_f;
    double n_pad;          // individuals of the tiles, padding included
};

// A fixed block of device memory that buffers are carved from, front to
// back, instead of being allocated (the allocation-free entry point).
struct Arena {
    char* base = nullptr;
    size_t cap = 0, off = 0;
};

// Grow-only device buffer: hipMalloc, or a piece of `arena` when set (then
// release() only forgets the piece; the arena is reset as a whole).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int device = -1;
    Arena* arena = nullptr;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Grow-only host staging buffer: pinned (hipHostMalloc), or pageable host
// memory when `pageable` (no HIP allocation; copies through it are staged
// by the runtime and return once the host side is consumed or filled).
struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool pageable = false;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Host-side band plan of the 64x64 tiles (I, J): per block row I its first
// column block j0[I] and its first tile tstart[I].
struct BandPlan {
    uint32_t m_pad = 0, nb = 0, ntiles = 0;
    std::vector<int> j0, tstart;
    uint64_t pairs_included = 0;
};

// Host arrays of a call, kept (with their capacity) from call to call so that
// a repeated call touches no freshly mapped pages.
struct HostScratch {
    std::vector<uint64_t> lo;
    BandPlan plan;
    std::vector<double> maf, aux;
    std::vector<double> coord;  // window coordinates of the pair-bound scan
    std::vector<uint8_t> round32;
    std::vector<uint64_t> chunk_end;  // ends of the reference's chunks (progress)
};

// One in-flight computation: its own stream, events and grow-only buffers.
struct Workspace {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // read-backs that overlap the band kernel
    hipStream_t up_stream = nullptr;    // annotation upload beside the tiling
    hipStream_t plan_stream = nullptr;  // band plan upload beside the tiling
    hipEvent_t plan_ev = nullptr;       // the band plan is on the device
    hipEvent_t ev[10] = {};
    DevBuf tiled, stats, bed_idx, lo, info, aeff, annot, pq, maf, mafd, tiles, j0, tstart,
        rowp, colp, l2, sk_cols, sk_norm, scratch;
    PinnedBuf pinned;       // SNP row indices (upload in flight while the host plans)
    PinnedBuf pinned_meta;  // pair bounds, band plan, stats read-back
    PinnedBuf pinned_maf;   // MAF read-back (during the band kernel)
    PinnedBuf pinned_flag;  // the sketch's integer-range flag (read before the band)
    // Large copies of caller memory (annotation upload, multi-annotation
    // results) pass through kRingSlots pinned slots of kRingSlot bytes: the
    // host's task pool fills or drains one slot while DMA moves another.
    static constexpr size_t kRingSlot = 8u << 20;
    static constexpr int kRingSlots = 4;
    PinnedBuf ring;
    hipEvent_t ring_ev[kRingSlots] = {};
    // a large result is reduced in up to kRedGroups groups of annotations,
    // each copied back as soon as it is done
    static constexpr uint32_t kRedGroups = 8;
    hipEvent_t red_ev[kRedGroups] = {};
    // with a progress callback the band runs in up to kProgPieces launches
    // of consecutive block rows, each followed by one of these events
    static constexpr uint32_t kProgPieces = 8;
    hipEvent_t prog_ev[kProgPieces] = {};
    HostScratch host;
    size_t bytes() const;  // device bytes held by the buffers
    hipError_t init(int dev);
    void trim();           // frees the device buffers (streams and events stay)
    void destroy();
    // Allocation-free mode: device buffers come from `a`, host staging is
    // pageable.  reset_arena() starts the next call's carving at the front.
    void use_arena(Arena* a);
    void reset_arena();
};

Workspace* acquire_workspace(int device);
void release_workspace(Workspace* ws);

}  // namespace ldsc

namespace ldsc {
// Device-side CountSketch plan for one sketch dimension d (see sketch.hip).
// Members are put in bucket order within chunks of `chunk` members (a
// multiple of the 128-member K-block): without --keep, chunks of
// kSkMembers, so that a chunk's members come from one short window of each
// .bed row; with --keep, one chunk of all members.
constexpr uint32_t kSkMembers = 16384;

// LDS slot of 16-byte unit u of a row quad in the chunk gather of engine.hip
// (member-bytes 4u .. 4u+3 of its four rows).  The staging writes of a wave
// cover units 4l + t (lane l, step t); XOR-ing the low two bits with bits
// 3-4 puts eight consecutive lanes in eight distinct groups of four banks.
// A bijection within aligned groups of eight units.
__host__ __device__ inline uint32_t gc_slot(uint32_t u) { return u ^ ((u >> 3) & 3u); }
// Where member `mem` of a chunk sits in such a staged quad: its LDS word
// (16-byte slot gc_slot(mem / 16), word (mem / 4) % 4) times 4, plus its
// 2-bit field within each byte (mem % 4).  Past the last member: the word
// after the window (kSkMembers / 4), which reads as code 01 in every byte.
__host__ __device__ inline uint16_t gc_code(uint32_t mem) {
    return (uint16_t)((((gc_slot(mem >> 4) << 2) | ((mem >> 2) & 3u)) << 2) | (mem & 3u));
}
constexpr uint16_t kGcPastCode = (uint16_t)((kSkMembers / 4) << 2);
struct SketchPlan {
    uint64_t d = 0, d_pad = 0, nblocks = 0, ntiles = 0, chunk = 0, nchunks = 0;
    int32_t* d_fam = nullptr;     // one chunk: FAM row of each position, lane-major map
    uint16_t* d_local = nullptr;  // chunks: gc_code of the member at each position (its
                                  // LDS word and field in the staged chunk window;
                                  // kGcPastCode past the last member), nkb * 128 entries
    void* d_ptiles = nullptr;     // ntiles x 1 KiB signed int8 MFMA A-operand tiles (k_sketch_fused)
    // k_sketch: e2m1 A-operand tiles (1 KiB each) of 64-position K-steps,
    // per bucket block its first tile and number of tiles, and each tile's
    // K-step
    void* d_ptiles4 = nullptr;
    int4* d_blocks = nullptr;
    int32_t* d_tile_ks = nullptr;
    int32_t* d_tsum = nullptr;    // per bucket (d_pad, zero padded): T_b = sum of its signs
    double tt = 0.0;              // sum over buckets of T_b^2
    // chunk-local order only: per chunk c, the chunk-local position where
    // bucket block B starts (nchunks x (nblocks + 1)), and the first P tile of
    // block B in chunk c (nblocks x nchunks)
    uint32_t* d_bo = nullptr;
    int32_t* d_bstart = nullptr;
    // chunk-local plans with at most kSdMaxBlocks bucket blocks:
    // per individual, its entry in the P operand of its pair of K-steps
    // (lane | word << 6 | K-step << 8 | block << 9 | nibble << 12 |
    // negative << 15), for k_sketch_dense
    uint16_t* d_pcode = nullptr;
    uint8_t* d_miss = nullptr;  // a row of all-missing codes (0x55), 32 ceil(N / 128) bytes
    void release();
};
}  // namespace ldsc

struct ldsc_panel {
    int device = 0;
    uint64_t n_fam = 0, m_bim = 0, bps = 0, n = 0;
    uint64_t stride = 0;         // bytes between resident rows: bps rounded up to 16
    uint8_t* d_slabs = nullptr;  // m_bim rows of the .bed body, `stride` apart, zero padded
    int64_t* d_keep = nullptr;   // n FAM rows, or nullptr (all individuals)
    int32_t* d_keep32 = nullptr; // the same rows as an int32 lane-major gather map
    std::vector<int64_t> keep;
    std::mutex sketch_mu;
    std::vector<ldsc::SketchPlan*> sketch_plans;
};

namespace ldsc {

// Tiles the call's SNP rows (d_bed_idx[j], or base + j when d_bed_idx is
// null) into ws->tiled and counts their integer stats
// into ws->stats.  `d_map` (nullable) gives, for each of the panel's N
// individuals in the order wanted, its FAM row (--keep order, or the
// CountSketch bucket order) in the lane-major layout of lane_major_map; it
// is applied by an LDS-staged row gather.  *parts receives the number of
// integer-stat slots written after ws->stats' first m_pad x 4 words.  With
// `tile` false and a map, only the gathered rows are produced (ws->scratch,
// ceil(N/16) dwords per SNP, SNP-major) with their stats in slot 0.
// `vcode` selects the dosage-value tile form of the e2m1 band kernel
// (engine.hip) instead of the PLINK-code form above.
int gather_and_tile(const ldsc_panel* panel, const uint64_t* d_bed_idx, uint64_t base,
                    uint64_t m, const int32_t* d_map, uint64_t m_pad, bool vcode, bool tile,
                    Workspace* ws, uint32_t* parts, char* err, uint64_t err_len);

// CountSketch plans with chunk-local bucket order (SketchPlan::d_local):
// tiles the call's SNP rows in that order straight from their .bed windows
// into ws->tiled; one integer-stat slot per chunk (*parts = chunks).
int gather_tile_chunks(const ldsc_panel* panel, const uint64_t* d_bed_idx, uint64_t base,
                       uint64_t m, const uint16_t* d_local, uint32_t chunk, uint64_t m_pad,
                       Workspace* ws, uint32_t* parts, char* err, uint64_t err_len);

// The exact / sketch engine on workspace `ws` (which must belong to
// panel->device).  Synchronous: outputs are written when it returns.
int run_l2(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
           const double* annot, uint64_t K, double* l2_out, double* maf_out,
           ldsc_l2_perf* perf, ldsc_progress_cb cb, void* cb_ctx, Workspace* ws, char* err,
           uint64_t err_len);

// run_l2 in three phases, so that the request server can put the band
// contractions of several calls into one launch:
//  * l2_begin: validation, tiling, host planning, uploads; everything the
//    band kernel reads is queued on ws->stream before ws->ev[3].  Sets
//    call->done when nothing is left to run (m == 0).
//  * l2_band: the band kernel of one call on ws->stream; or
//    band_batch_launch: the bands of several calls of one form and device
//    in one launch on stream `bst`, after each call's ev[3]; it records each
//    call's ws->ev[4] after the launch.
//  * l2_end: waits for ev[4] on ws->stream, reduces, copies the results
//    back; synchronous.
enum BandForm { kBandFp4 = 0, kBandI8 = 1, kBandF32 = 2, kBandSketch = 3, kBandSketch8 = 4 };
inline bool is_sketch_form(int form) { return form == kBandSketch || form == kBandSketch8; }
struct SketchPlan;
struct L2Call {
    const ldsc_panel* panel = nullptr;
    const ldsc_snp* snps = nullptr;
    uint64_t m = 0, K = 1, d = 0;
    const ldsc_l2_config* cfg = nullptr;
    double* l2_out = nullptr;
    double* maf_out = nullptr;
    ldsc_l2_perf* perf = nullptr;
    ldsc_progress_cb cb = nullptr;
    void* cb_ctx = nullptr;
    Workspace* ws = nullptr;
    double t_start = 0, t_prep = 0;
    uint32_t m_pad = 0;
    int form = kBandFp4;
    bool use_f32 = false, early_maf = false, early_stats = false, stage_l2 = false;
    bool stats_fetched = false;  // the integer stats come back on the copy stream (ev[7])
    bool any_full = false, done = true;
    size_t off_st = 0, off_l2 = 0;  // in ws->pinned_meta
    SketchPlan* sketch_plan = nullptr;
    BandArgs ba{};
    // A large annotation goes up in row segments; the band of segment s
    // (row blocks seg_I[s] .. seg_I[s+1]-1) starts once rows up to
    // 64 seg_I[s+1] are on the device, while the next segment uploads.
    static constexpr uint32_t kMaxSeg = 4;
    const double* annot = nullptr;
    uint32_t nseg = 1, Kp = 1;
    uint32_t seg_I[kMaxSeg + 1] = {};
    uint32_t seg_dl_a = 0, seg_dl_b = 0;  // reduced rows l2_end brings back
    bool has_pq = false;
    // Progress points of the band (callback given): once prog_ev[p] has
    // fired, every pair of the SNPs below prog_snps[p] has been evaluated.
    static constexpr uint32_t kMaxProg = 8;
    uint32_t nprog = 0;
    uint32_t prog_snps[kMaxProg] = {};
    hipEvent_t prog_ev[kMaxProg] = {};
    bool sk_fused = false;        // bucket sums by the one-pass kernel (no tiled rows)
    bool bed_contiguous = false;  // SNP rows base .. base + m - 1 (no index upload)
};
// Argument table of one batched band launch (device copy + pinned staging),
// reused once the launch that read it has finished.
struct BandBatchSlot {
    static constexpr uint32_t kMaxCalls = 64;
    static constexpr size_t kBytes = kMaxCalls * sizeof(BandArgs) + (kMaxCalls + 1) * 4;
    void* d_args = nullptr;
    void* h_args = nullptr;
    hipEvent_t done = nullptr;
    bool used = false;
    void release();
};
// segment_annot: a large annotation may go up in row segments that l2_band
// interleaves with the band (one-call path only; a batched band needs the
// whole annotation at l2_begin's end).
int l2_begin(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
             const double* annot, uint64_t K, double* l2_out, double* maf_out, ldsc_l2_perf* perf,
             ldsc_progress_cb cb, void* cb_ctx, Workspace* ws, L2Call* call, char* err,
             uint64_t err_len, bool segment_annot = false);
int l2_band(L2Call* call, char* err, uint64_t err_len);
int band_batch_launch(L2Call* const* calls, uint32_t ncalls, BandBatchSlot* slot, hipStream_t bst,
                      char* err, uint64_t err_len);
int l2_end(L2Call* call, char* err, uint64_t err_len);

// Tiles of the call's band: the pair bounds of its SNP list (into `lo`) and
// the 64x64 tiles (I, J), lo[64 I] / 64 <= J <= I, they span.
uint64_t band_tile_count(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                         std::vector<uint64_t>& lo);
// Device bytes a call of this shape allocates in a fresh workspace (upper
// bound, for admission control); `tiles` from band_tile_count.
size_t l2_workspace_bytes(const ldsc_panel* panel, uint64_t m, const ldsc_l2_config* cfg,
                          uint64_t K, uint64_t tiles, bool annotated);

}  // namespace ldsc
