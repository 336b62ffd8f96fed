// engine.hip -- resident panels and the exact LD-score engine for gfx950.
//
// Pipeline of one ldsc_l2_compute call (exact mode):
//   1. k_tile_stats   gather the call's SNP slabs (bed_idx order, --keep
//                     subset) from the resident .bed copy, write them in the
//                     MFMA tile layout (device.h) and count per-SNP integer
//                     stats (sum, count, sum_sq, het) -- HBM-bound.
//   2. k_snp_info     per-SNP constants of the finish: exact-integer (C, S,
//                     1/sqrt(C*Q)) or the reference's f32 standardization LUT.
//   3. k_band_fp4 / k_band_i8 / k_band_f32_2w
//                     the 64x64 tiles of the SNP band.  Exact form: four
//                     exact integer products G.G, G.M, M.G, M.M on the e2m1
//                     MFMA (one wave per tile) or v_mfma_i32_32x32x32_i8
//                     (two waves per tile), FP64 finish of the mean-imputed
//                     covariance and r^2.  f32 form: the reference's
//                     standardized values on v_mfma_f32_32x32x2f32 (two
//                     waves per tile).  Then the unbiased r^2, the pair
//                     bound mask and the r2u . A_eff contractions for both
//                     members of each pair, written as per-tile partial sums
//                     (no atomics).
//   4. k_reduce       per-SNP sum of its tile partials in a fixed order plus
//                     the self term -> deterministic l2.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <type_traits>

#include "band_common.h"

namespace ldsc {

#define HIPC(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(err, err_len, e_ == hipErrorOutOfMemory ? LDSC_ERR_OOM            \
                                                                 : LDSC_ERR_DEVICE,       \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                    \
    } while (0)


// ------------------------------------------------------------------------
// Buffers and workspaces
// ------------------------------------------------------------------------
hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (arena) {
        const size_t want = round_up(std::max<size_t>(bytes, 256), 256);
        if (arena->off + want > arena->cap) return hipErrorOutOfMemory;
        p = arena->base + arena->off;
        arena->off += want;
        cap = want;
        return hipSuccess;
    }
    release();
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        return e;
    }
    cap = want;
    return hipSuccess;
}
void DevBuf::release() {
    if (p && !arena) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

void BandBatchSlot::release() {
    if (d_args) (void)hipFree(d_args);
    if (h_args) (void)hipHostFree(h_args);
    if (done) (void)hipEventDestroy(done);
    d_args = h_args = nullptr;
    done = nullptr;
    used = false;
}

hipError_t PinnedBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    release();
    const size_t want = std::max<size_t>(bytes, 4096);
    if (pageable) {
        p = std::aligned_alloc(4096, round_up(want, 4096));
        if (!p) return hipErrorOutOfMemory;
        cap = want;
        return hipSuccess;
    }
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        cap = 0;
        return e;
    }
    cap = want;
    return hipSuccess;
}
void PinnedBuf::release() {
    if (p && pageable)
        std::free(p);
    else if (p)
        (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

size_t Workspace::bytes() const {
    const DevBuf* all[] = {&tiled, &stats, &bed_idx, &lo, &info, &aeff, &annot, &pq, &maf, &mafd,
                           &tiles, &j0, &tstart, &rowp, &colp, &l2, &sk_cols, &sk_norm, &scratch};
    size_t s = 0;
    for (auto* b : all) s += b->cap;
    return s;
}

hipError_t Workspace::init(int dev) {
    device = dev;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return e;
    e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamCreateWithFlags(&up_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamCreateWithFlags(&plan_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipEventCreateWithFlags(&plan_ev, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    for (auto& x : ev) {
        e = hipEventCreate(&x);
        if (e != hipSuccess) return e;
    }
    for (auto& x : ring_ev) {
        e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    for (auto& x : red_ev) {
        e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    for (auto& x : prog_ev) {
        e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

void Workspace::trim() {
    DevBuf* all[] = {&tiled, &stats, &bed_idx, &lo, &info, &aeff, &annot, &pq, &maf, &mafd,
                     &tiles, &j0, &tstart, &rowp, &colp, &l2, &sk_cols, &sk_norm, &scratch};
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    if (up_stream) (void)hipStreamSynchronize(up_stream);
    if (plan_stream) (void)hipStreamSynchronize(plan_stream);
    for (auto* b : all) b->release();
}

void Workspace::use_arena(Arena* a) {
    DevBuf* all[] = {&tiled, &stats, &bed_idx, &lo, &info, &aeff, &annot, &pq, &maf, &mafd,
                     &tiles, &j0, &tstart, &rowp, &colp, &l2, &sk_cols, &sk_norm, &scratch};
    for (auto* b : all) b->arena = a;
    PinnedBuf* host[] = {&pinned, &pinned_meta, &pinned_maf, &pinned_flag, &ring};
    for (auto* b : host) b->pageable = true;
}

void Workspace::reset_arena() {
    DevBuf* all[] = {&tiled, &stats, &bed_idx, &lo, &info, &aeff, &annot, &pq, &maf, &mafd,
                     &tiles, &j0, &tstart, &rowp, &colp, &l2, &sk_cols, &sk_norm, &scratch};
    for (auto* b : all) {
        b->p = nullptr;
        b->cap = 0;
        if (b->arena) b->arena->off = 0;
    }
}

void Workspace::destroy() {
    trim();
    pinned.release();
    pinned_meta.release();
    pinned_maf.release();
    pinned_flag.release();
    ring.release();
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    for (auto& x : ring_ev)
        if (x) (void)hipEventDestroy(x);
    for (auto& x : red_ev)
        if (x) (void)hipEventDestroy(x);
    for (auto& x : prog_ev)
        if (x) (void)hipEventDestroy(x);
    if (plan_ev) (void)hipEventDestroy(plan_ev);
    if (stream) (void)hipStreamDestroy(stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (up_stream) (void)hipStreamDestroy(up_stream);
    if (plan_stream) (void)hipStreamDestroy(plan_stream);
    copy_stream = nullptr;
    up_stream = nullptr;
    plan_stream = nullptr;
    plan_ev = nullptr;
    stream = nullptr;
}

// Host -> device copy of `bytes` of caller memory through the pinned ring,
// on stream s, filled by the side copy pool (the caller runs it beside the
// host's planning).  With `fill`, slot bytes [off, off+len) of the
// destination come from fill(slot, off, len) instead of src + off.  A slot
// is refilled only once its previous DMA (of any earlier transfer) is done.
// Returns once the last slot is filled; its DMA may still be in flight
// (later work on s is ordered after it).
using RingFill = std::function<void(char* slot, size_t off, size_t len)>;
static hipError_t ring_upload(Workspace* ws, void* dst, const void* src, size_t bytes,
                              hipStream_t s, const RingFill* fill = nullptr) {
    constexpr size_t SB = Workspace::kRingSlot;
    constexpr int NS = Workspace::kRingSlots;
    if (bytes <= SB && !fill)  // small: the runtime's own staging
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    hipError_t e = ws->ring.ensure(SB * NS);
    if (e != hipSuccess) return e;
    char* ring = ws->ring.as<char>();
    size_t c = 0;
    for (size_t off = 0; off < bytes; off += SB, ++c) {
        const int slot = (int)(c % NS);
        const size_t len = std::min(SB, bytes - off);
        if ((e = hipEventSynchronize(ws->ring_ev[slot])) != hipSuccess) return e;
        if (fill)
            (*fill)(ring + slot * SB, off, len);
        else
            parallel_copy(ring + slot * SB, static_cast<const char*>(src) + off, len, /*side=*/true);
        e = hipMemcpyAsync(static_cast<char*>(dst) + off, ring + slot * SB, len,
                           hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        if ((e = hipEventRecord(ws->ring_ev[slot], s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Elements [e0, e1) of a block holding rows [a, a + rows) of every column
// of a column-major matrix with m rows (the block column-major with column
// stride `rows`), copied from the matrix into the block or back, in column
// runs split over the main or side pool.
static void copy_block_rows(char* block, size_t e_base, double* mat, uint64_t m, uint64_t a,
                            uint64_t rows, size_t e0, size_t e1, bool to_block, bool side) {
    const int T = copy_task_count((e1 - e0) * 8, side);
    auto task = [&](int t) {
        size_t e = e0 + (e1 - e0) * t / T;
        const size_t end = e0 + (e1 - e0) * (t + 1) / T;
        while (e < end) {
            const uint64_t k = e / rows, row = e % rows;
            const size_t run = std::min<size_t>(end - e, rows - row);
            char* b = block + (e - e_base) * 8;
            double* x = mat + k * m + a + row;
            if (to_block)
                std::memcpy(b, x, run * 8);
            else
                std::memcpy(x, b, run * 8);
            e += run;
        }
    };
    if (side)
        parallel_tasks_side(T, task);
    else
        parallel_tasks(T, task);
}

// Rows [r0, r1) of every column of the column-major m x K annotation go up
// as one block, column-major with column stride r1 - r0, at row offset r0
// of the device copy (segments lie back to back in row order).
static hipError_t upload_annot_rows(Workspace* ws, const double* annot, uint64_t m, uint64_t K,
                                    uint64_t r0, uint64_t r1, hipStream_t s) {
    const uint64_t rows = r1 - r0;
    const RingFill fill = [=](char* slot, size_t off, size_t len) {
        copy_block_rows(slot, off / 8, const_cast<double*>(annot), m, r0, rows, off / 8,
                        (off + len) / 8, /*to_block=*/true, /*side=*/true);
    };
    return ring_upload(ws, ws->annot.as<double>() + r0 * K, nullptr, rows * K * 8, s, &fill);
}

// Device -> host copy into caller memory through the pinned ring on stream
// s; ready(end) is called before the copy of bytes [.., end) is queued (to
// order it after their producer).  With `drain`, slot bytes [off, off+len)
// go out through drain(slot, off, len) instead of to dst + off.  Returns
// when every byte has landed.
using RingDrain = std::function<void(const char* slot, size_t off, size_t len)>;
static hipError_t ring_download(Workspace* ws, void* dst, const void* src, size_t bytes,
                                hipStream_t s, const std::function<hipError_t(size_t)>& ready,
                                const RingDrain* drain = nullptr) {
    constexpr size_t SB = Workspace::kRingSlot;
    constexpr int NS = Workspace::kRingSlots;
    hipError_t e = ws->ring.ensure(SB * NS);
    if (e != hipSuccess) return e;
    char* ring = ws->ring.as<char>();
    const size_t nc = (bytes + SB - 1) / SB;
    auto issue = [&](size_t c) {
        const int slot = (int)(c % NS);
        const size_t off = c * SB, len = std::min(SB, bytes - off);
        hipError_t r = ready(off + len);
        // after any upload DMA still reading the slot
        if (r == hipSuccess) r = hipStreamWaitEvent(s, ws->ring_ev[slot], 0);
        if (r == hipSuccess)
            r = hipMemcpyAsync(ring + slot * SB, static_cast<const char*>(src) + off, len,
                               hipMemcpyDeviceToHost, s);
        return r != hipSuccess ? r : hipEventRecord(ws->ring_ev[slot], s);
    };
    for (size_t c = 0; c < nc && c < (size_t)NS; ++c)
        if ((e = issue(c)) != hipSuccess) return e;
    for (size_t c = 0; c < nc; ++c) {
        const int slot = (int)(c % NS);
        const size_t off = c * SB, len = std::min(SB, bytes - off);
        if ((e = hipEventSynchronize(ws->ring_ev[slot])) != hipSuccess) return e;
        if (drain)
            (*drain)(ring + slot * SB, off, len);
        else
            parallel_copy(static_cast<char*>(dst) + off, ring + slot * SB, len);
        if (c + NS < nc && (e = issue(c + NS)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// LD scores of rows [a, b), reduced as one block (column-major within it,
// at row offset a of the device result), into the column-major m x K `out`.
static hipError_t download_l2_rows(Workspace* ws, double* out, uint64_t m, uint64_t K, uint64_t a,
                                   uint64_t b, hipStream_t s) {
    const uint64_t rows = b - a;
    const RingDrain drain = [=](const char* slot, size_t off, size_t len) {
        copy_block_rows(const_cast<char*>(slot), off / 8, out, m, a, rows, off / 8,
                        (off + len) / 8, /*to_block=*/false, /*side=*/false);
    };
    return ring_download(ws, nullptr, ws->l2.as<double>() + a * K, rows * K * 8, s,
                         [](size_t) { return hipSuccess; }, &drain);
}

static std::mutex g_ws_mu;
static std::vector<Workspace*> g_ws_free;

Workspace* acquire_workspace(int device) {
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        for (size_t i = 0; i < g_ws_free.size(); ++i) {
            if (g_ws_free[i]->device == device) {
                Workspace* w = g_ws_free[i];
                g_ws_free.erase(g_ws_free.begin() + i);
                return w;
            }
        }
    }
    Workspace* w = new Workspace();
    if (w->init(device) != hipSuccess) {
        w->destroy();
        delete w;
        return nullptr;
    }
    return w;
}

void release_workspace(Workspace* w) {
    if (!w) return;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws_free.push_back(w);
}

// ------------------------------------------------------------------------
// Kernel 1a: row gather in a chosen individual order (--keep, bucket order)
// ------------------------------------------------------------------------
// R SNP rows at a time are staged in LDS (coalesced loads).  Each thread owns
// output dword q (members 16q .. 16q+15): it loads its 16 source rows from the
// lane-major map once (16 coalesced loads) and assembles that dword for all R
// rows from random LDS reads.  Output rows are `bps_out` = 4 * n16 bytes;
// members at or beyond N get code 01 (missing).
template <int R>
__global__ __launch_bounds__(512) void k_permute_rows(const uint8_t* __restrict__ slabs,
                                                       uint64_t pitch, uint32_t dwords,
                                                       uint32_t ldsrow,
                                                       const uint64_t* __restrict__ bed_idx,
                                                       uint64_t base, uint32_t m,
                                                       const int32_t* __restrict__ map_t,
                                                       uint32_t n16, uint8_t* __restrict__ out) {
    extern __shared__ uint8_t rows[];
    const int tid = threadIdx.x;
    for (uint32_t j0 = blockIdx.x * R; j0 < m; j0 += gridDim.x * R) {
        const uint32_t nr = m - j0 < (uint32_t)R ? m - j0 : (uint32_t)R;
        __syncthreads();
        for (uint32_t r = 0; r < nr; ++r) {
            const uint64_t row = bed_idx ? bed_idx[j0 + r] : base + j0 + r;
            const uint32_t* s4 = reinterpret_cast<const uint32_t*>(slabs + row * pitch);
            uint32_t* d4 = reinterpret_cast<uint32_t*>(rows + (size_t)r * ldsrow);
            for (uint32_t q = tid; q < dwords; q += blockDim.x) d4[q] = s4[q];
        }
        __syncthreads();
        for (uint32_t q = tid; q < n16; q += blockDim.x) {
            int32_t sr[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) sr[e] = map_t[(size_t)e * n16 + q];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if ((uint32_t)r < nr) {
                    const uint8_t* row = rows + (size_t)r * ldsrow;
                    uint32_t w = 0;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const uint32_t code =
                            sr[e] >= 0 ? (row[sr[e] >> 2] >> (2 * (sr[e] & 3))) & 3u : 1u;
                        w |= code << (2 * e);
                    }
                    reinterpret_cast<uint32_t*>(out + (uint64_t)(j0 + r) * n16 * 4)[q] = w;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// Kernel 1b: tile + integer stats
// ------------------------------------------------------------------------
// Workgroup (g, c) owns SNP group g (32 rows) and K-blocks
// [c * kTsChunk, (c + 1) * kTsChunk).  Each pass stages 8 K-blocks of the 32
// rows in LDS: wave w reads rows 8w .. 8w+7 with one dword per lane (256
// contiguous bytes of a row per load), then every thread transposes the
// dword it owns in each of the 8 output tiles and stores it (each wave
// stores two whole 128-byte runs of a tile).  Integer stats of the chunk go
// to their own slot; k_snp_info adds the chunks in a fixed order.  Rows are
// `pitch` bytes apart and dword-aligned; dwords past the row's last
// individual read as code 01 (missing).
constexpr int kTsPass = 8;                 // K-blocks per LDS pass (64 dwords of a row)
constexpr int kTsChunk = 8 * kTsPass;      // K-blocks per workgroup
constexpr int kTsStride = 72;              // LDS row stride in dwords, 8 mod 32:
                                           // the transposing reads hit 32 banks

// 4x4 transpose of 2-bit cells: individual 4f+k moves from bit 8f+2k to bit
// 8k+2f (two delta swaps).
__device__ __forceinline__ uint32_t transpose_cells(uint32_t x) {
    uint32_t d = ((x >> 6) ^ x) & 0x00CC00CCu;
    x ^= d ^ (d << 6);
    d = ((x >> 12) ^ x) & 0x0000F0F0u;
    return x ^ d ^ (d << 12);
}

// Number of K-block chunks (integer stat slots) of a tiling over n individuals.
static inline uint32_t tile_parts(uint64_t n) {
    const uint64_t nkb = round_up(n, kKBlock) / kKBlock;
    return (uint32_t)((nkb + kTsChunk - 1) / kTsChunk);
}

// Dosage-value code of 16 cells: PLINK code 00 (dosage 2) -> 2, 10
// (dosage 1) -> 1, 11 (dosage 0) -> 0, 01 (missing) -> 3, i.e. bit 0 =
// L xor H, bit 1 = not H.
__device__ __forceinline__ uint32_t value_codes(uint32_t raw) {
    const uint32_t L = raw & 0x55555555u, H = (raw >> 1) & 0x55555555u;
    return (L ^ H) | ((H ^ 0x55555555u) << 1);
}

// VCODE = false: PLINK-code tiles (device.h), for the int8, FP32 and
//   CountSketch kernels.
// VCODE = true: dosage-value tiles of the e2m1 kernel, also 1 KiB per group
//   and K-block: lane l (SNP l%32) holds 4 dwords, dword 2u + i covering
//   individuals u*64 + (l/32)*32 + 16i + f in its 2-bit fields f (value
//   codes, missing and padding = 3), i.e. row dword d = 2t + h of a K-block
//   lands unchanged but re-coded in lane s + 32(t%2), dword 2(t/2) + h.
template <bool VCODE>
__global__ __launch_bounds__(256) void k_tile_stats(const uint8_t* __restrict__ slabs,
                                                     uint64_t pitch, uint32_t n,
                                                     const uint64_t* __restrict__ bed_idx,
                                                     uint64_t base, uint32_t m, uint32_t nkb,
                                                     uint32_t m_pad,
                                                     uint32_t* __restrict__ tiled,
                                                     uint32_t* __restrict__ part) {
    __shared__ uint32_t lds[kGroup * kTsStride];
    const int tid = threadIdx.x, wave = tid >> 6, lane_ld = tid & 63;
    const uint32_t g = blockIdx.x, chunk = blockIdx.y;
    const uint32_t kb_begin = chunk * kTsChunk;
    const uint32_t kb_end = min(nkb, kb_begin + kTsChunk);
    // load side: 8 rows per wave; loads are clamped to the row's last dword
    // (in bounds) and masked when staged
    const uint32_t* rowp[8];
    uint32_t nv[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint32_t snp = g * kGroup + 8 * wave + r;
        const uint64_t src = snp < m ? (bed_idx ? bed_idx[snp] : base + snp) : base;
        rowp[r] = reinterpret_cast<const uint32_t*>(slabs + src * pitch);
        nv[r] = snp < m ? n : 0;
    }
    const uint64_t qlast = (n - 1) / 16;
    // store side: thread -> (SNP s, dword t, half h) of every tile; a
    // half-wave holds 4 SNPs x 4 t x 2 h, so its LDS reads s*72 + 2t + h
    // fall in 32 distinct banks
    const int h = tid & 1, t = (tid >> 1) & 3;
    const int s = 8 * wave + 4 * ((tid >> 5) & 1) + ((tid >> 3) & 3);
    const int lane = s + 32 * h;
    uint32_t* out = VCODE ? tiled + ((size_t)g * nkb * 64 + s + 32 * (t & 1)) * 4 + 2 * (t >> 1) + h
                          : tiled + ((size_t)g * nkb * 64 + lane) * 4 + t;
    constexpr int kTileDw = 256;
    const uint32_t* src_lds = lds + s * kTsStride + 2 * t + h;
    uint32_t n2 = 0, n1 = 0, n0 = 0;
    uint32_t v[8];
    uint64_t q = (uint64_t)kb_begin * 8 + lane_ld;
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = rowp[r][q < qlast ? q : qlast];
    for (uint32_t kb0 = kb_begin; kb0 < kb_end; kb0 += kTsPass, q += 8 * kTsPass) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) lds[(8 * wave + r) * kTsStride + lane_ld] = ts_mask(v[r], q, nv[r]);
        __syncthreads();
        if (kb0 + kTsPass < kb_end) {
            const uint64_t qn = q + 8 * kTsPass;
#pragma unroll
            for (int r = 0; r < 8; ++r) v[r] = rowp[r][qn < qlast ? qn : qlast];
        }
        uint32_t x[kTsPass];
#pragma unroll
        for (int k = 0; k < kTsPass; ++k) {
            const uint32_t raw = src_lds[8 * k];
            const uint32_t L = raw & 0x55555555u, H = (raw >> 1) & 0x55555555u;
            n2 += __popc(~(L | H) & 0x55555555u);  // 00: dosage 2
            n1 += __popc(~L & H);                   // 10: dosage 1 (het)
            n0 += __popc(L & H);                    // 11: dosage 0
            x[k] = VCODE ? value_codes(raw) : transpose_cells(raw);
        }
        // K-blocks past nkb (last pass) hold code 01 only: counted as
        // nothing above, and not stored
        uint32_t* o = out + (size_t)kb0 * kTileDw;
        if (kb0 + kTsPass <= kb_end) {
#pragma unroll
            for (int k = 0; k < kTsPass; ++k) o[k * kTileDw] = x[k];
        } else {
#pragma unroll
            for (int k = 0; k < kTsPass; ++k)
                if (kb0 + k < kb_end) o[k * kTileDw] = x[k];
        }
    }
    // the 8 threads of a SNP are consecutive: fold them, then one writes
    uint32_t c[4] = {2 * n2 + n1, n2 + n1 + n0, 4 * n2 + n1, n1};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        c[q] += __shfl_xor(c[q], 1);
        c[q] += __shfl_xor(c[q], 2);
        c[q] += __shfl_xor(c[q], 4);
    }
    if ((tid & 7) == 0) {
        uint4* dst = reinterpret_cast<uint4*>(part) + (size_t)chunk * m_pad + g * kGroup + s;
        *dst = make_uint4(c[0], c[1], c[2], c[3]);
    }
}

// ------------------------------------------------------------------------
// Kernel 1c: four-row gather through byte-interleaved LDS
// ------------------------------------------------------------------------
// For --keep subsets and the CountSketch bucket order: four SNP rows per
// pass are staged in LDS byte-interleaved (word k holds byte k of the four
// rows), so one 4-byte LDS read fetches a member's code in all four rows.
// Thread q builds output dword q (members 16q .. 16q+15, map_t lane-major,
// -1 past the last member -> code 01) of the four rows, with the next
// dword's 16 map entries already loading, and stores the four dwords into
// the gathered rows (`n16` dwords each, coalesced across the wave).  The
// rows' integer stats (sum, count, sum_sq, het) go to part[j] as uint4.

// Output forms of the gather: the gathered rows (n16 dwords each), or
// straight into the tile layouts of k_tile_stats (PLINK-code tiles, or
// dosage-value tiles), which skips writing and re-reading the rows.
enum GatherForm { kGatherRows = 0, kGatherCodeTiles = 1, kGatherValueTiles = 2 };

// Tile dword of gathered dword q (members 16q .. 16q+15) of SNP j, in the
// layouts k_tile_stats writes (nkb K-blocks per SNP group).
template <int FORM>
__device__ __forceinline__ size_t tile_dword(uint32_t j, uint32_t q, uint32_t nkb) {
    const uint32_t g = j / kGroup, s = j % kGroup, kb = q / 8, dq = q % 8;
    const size_t base = ((size_t)g * nkb + kb) * 64;
    if (FORM == kGatherCodeTiles)  // row dword 2t + h -> lane s + 32h, dword t
        return (base + s + 32 * (dq & 1)) * 4 + (dq >> 1);
    // row dword 2t + h -> lane s + 32(t%2), dword 2(t/2) + h
    return (base + s + 32 * ((dq >> 1) & 1)) * 4 + 2 * (dq >> 2) + (dq & 1);
}

// PF > 0: each thread keeps its PF source dwords per row of the next pass
// in registers, loaded while the current pass gathers (src_dw <= 512 PF).
// FORM != kGatherRows: `out` is the tile array; n_out = nkb * 8 dwords per
// SNP are written (members past the last read as code 01, missing), and
// padding SNPs j in [m, m_pad) are written as all missing.
template <int PF, int FORM>
__global__ __launch_bounds__(512) void k_gather4(const uint8_t* __restrict__ slabs, uint64_t pitch,
                                                  uint32_t src_dw,
                                                  const uint64_t* __restrict__ bed_idx,
                                                  uint64_t base, uint32_t m, uint32_t m_pad,
                                                  const int32_t* __restrict__ map_t, uint32_t n16,
                                                  uint32_t n_out, uint32_t* __restrict__ out,
                                                  uint32_t* __restrict__ part) {
    extern __shared__ uint32_t lds[];  // 4 * src_dw words
    __shared__ uint32_t red[8][12];
    const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63;
    auto row_ptr = [&](uint32_t j) {
        const uint64_t row = j < m ? (bed_idx ? bed_idx[j] : base + j) : 0;
        return reinterpret_cast<const uint32_t*>(slabs + row * pitch);
    };
    uint32_t pre[PF > 0 ? PF : 1][4];
    auto fetch = [&](uint32_t j0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t* rp = row_ptr(j0 + r);
            const bool v = j0 + r < m;
#pragma unroll
            for (int i = 0; i < PF; ++i) {
                const uint32_t q = tid + 512 * i;
                pre[i][r] = (v && q < src_dw) ? rp[q] : 0u;
            }
        }
    };
    if (PF > 0 && 4 * blockIdx.x < m_pad) fetch(4 * blockIdx.x);
    for (uint32_t rg = blockIdx.x; 4 * rg < m_pad; rg += gridDim.x) {
        const uint32_t j0 = 4 * rg;
        bool valid[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) valid[r] = j0 + r < m;
        __syncthreads();  // the previous pass is done with lds
        if (PF > 0) {
#pragma unroll
            for (int i = 0; i < PF; ++i) {
                const uint32_t q = tid + 512 * i;
                if (q < src_dw) {
                    uint32_t o[4];
                    transpose_bytes4(pre[i][0], pre[i][1], pre[i][2], pre[i][3], o);
                    reinterpret_cast<uint4*>(lds)[q] = make_uint4(o[0], o[1], o[2], o[3]);
                }
            }
        } else {
            const uint32_t* rp[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) rp[r] = row_ptr(j0 + r);
            for (uint32_t q = tid; q < src_dw; q += 512) {
                uint32_t o[4];
                transpose_bytes4(valid[0] ? rp[0][q] : 0u, valid[1] ? rp[1][q] : 0u,
                                 valid[2] ? rp[2][q] : 0u, valid[3] ? rp[3][q] : 0u, o);
                reinterpret_cast<uint4*>(lds)[q] = make_uint4(o[0], o[1], o[2], o[3]);
            }
        }
        __syncthreads();
        // the next pass's rows load while this one gathers
        if (PF > 0 && 4 * (rg + gridDim.x) < m_pad) fetch(4 * (rg + gridDim.x));
        uint32_t cnt[4][3];
#pragma unroll
        for (int r = 0; r < 4; ++r) cnt[r][0] = cnt[r][1] = cnt[r][2] = 0;
        int32_t sr[16];
        uint32_t q = tid;
#pragma unroll
        for (int e = 0; e < 16; ++e) sr[e] = q < n16 ? map_t[(size_t)e * n16 + q] : -1;
        for (; q < n_out; q += 512) {
            int32_t cur[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) cur[e] = sr[e];
#pragma unroll
            for (int e = 0; e < 16; ++e)
                sr[e] = q + 512 < n16 ? map_t[(size_t)e * n16 + q + 512] : -1;
            uint32_t acc[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int32_t s = cur[e];
                uint32_t c4 = 0x01010101u;  // missing in all four rows
                if (s >= 0) c4 = (lds[s >> 2] >> (2 * (s & 3))) & 0x03030303u;
                // code tiles: member e = 4f + k straight to the tiles' cell
                // order, bits 8k + 2f, through acc[k] bits 2f
                if (FORM == kGatherCodeTiles)
                    acc[e & 3] |= c4 << (2 * (e >> 2));
                else
                    acc[e >> 2] |= c4 << (2 * (e & 3));
            }
            uint32_t o[4];
            transpose_bytes4(acc[0], acc[1], acc[2], acc[3], o);  // o[r] = row r's dword
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t L = o[r] & 0x55555555u, H = (o[r] >> 1) & 0x55555555u;
                cnt[r][0] += __popc(~(L | H) & 0x55555555u);  // 00: dosage 2
                cnt[r][1] += __popc(~L & H);                   // 10: dosage 1
                cnt[r][2] += __popc(L & H);                    // 11: dosage 0
                if (FORM == kGatherRows) {
                    if (valid[r]) out[(size_t)(j0 + r) * n16 + q] = o[r];
                } else {
                    const uint32_t w = valid[r] ? o[r] : 0x55555555u;  // padding SNP: missing
                    out[tile_dword<FORM>(j0 + r, q, n_out / 8)] =
                        FORM == kGatherCodeTiles ? w : value_codes(w);
                }
            }
        }
        // block sums of the 4 x 3 counts: waves first, then across waves
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                uint32_t v = cnt[r][c];
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
                cnt[r][c] = v;
            }
        if (ln == 0)
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) red[wv][3 * r + c] = cnt[r][c];
        __syncthreads();
        if (tid < 4) {
            uint32_t n2 = 0, n1 = 0, n0 = 0;
            for (int w = 0; w < 8; ++w) {
                n2 += red[w][3 * tid + 0];
                n1 += red[w][3 * tid + 1];
                n0 += red[w][3 * tid + 2];
            }
            const bool v = j0 + tid < m;  // padding rows: nothing
            reinterpret_cast<uint4*>(part)[j0 + tid] =
                v ? make_uint4(2 * n2 + n1, n2 + n1 + n0, 4 * n2 + n1, n1) : make_uint4(0, 0, 0, 0);
        }
    }
}

// ------------------------------------------------------------------------
// Kernel 1d: chunk-local gather into dosage-value tiles (CountSketch order)
// ------------------------------------------------------------------------
// Workgroup h owns SNP rows 16h .. 16h+15 and walks the member chunks c of
// MEMBERS members (FAM rows c*MEMBERS ..): one MEMBERS/4-byte window of
// each .bed row per chunk.  Eight waves: waves w and w + 4 load rows
// 16h + 4w .. +3 (row quad w), each half of the window (1 KiB of one row
// per load), re-code them to dosage values, count their integer stats (the
// chunk's slot) and stage them byte-interleaved in LDS (word k of quad w =
// byte k of its four rows, in 16-byte slots placed by gc_slot).  Then
// thread t builds output dwords t + 512u of the chunk (16 positions each,
// local[] gives each position's member within the chunk) for the four
// quads and stores them into the dosage-value tiles of k_tile_stats<true>;
// positions past the last member and padding SNPs read as missing (value
// 3).  The next chunk's rows load into registers while this chunk's
// gather runs.  Only a pass that reaches past the last individual, or a
// quad with padding SNPs, masks its cells (a uniform branch).
constexpr int kGcRows = 16;
constexpr int kGcThreads = 512;

// Sum of v over the wave's 64 lanes (DPP within rows of 16, then the four
// rows' first lanes).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);   // quad_perm 1,0,3,2
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);   // quad_perm 2,3,0,1
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true);  // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true);  // row_mirror
    return __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) +
           __builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48);
}

template <uint32_t MEMBERS>
__global__ __launch_bounds__(kGcThreads) void k_gather_chunk(
    const uint8_t* __restrict__ slabs, uint64_t pitch, uint32_t src_dw, uint32_t n,
    const uint64_t* __restrict__ bed_idx, uint64_t base, uint32_t m, uint32_t m_pad,
    const uint16_t* __restrict__ local, uint32_t nkb, uint32_t nchunks,
    uint32_t* __restrict__ tiled, uint32_t* __restrict__ part) {
    constexpr uint32_t kGcWords = MEMBERS / 4;    // LDS words per quad
    constexpr uint32_t kGcStride = kGcWords + 4;  // quad stride; word kGcWords (its slot
                                                  // maps to itself) reads as value 3
                                                  // (missing) in every field
    constexpr uint32_t cw = MEMBERS / 16;         // window dwords per row
    constexpr int kPasses = (int)(cw / 256);      // 256 window dwords per pass
    constexpr int kHalf = kPasses / 2;            // passes of one wave
    static_assert(kPasses % 2 == 0, "the two waves of a quad take half the passes each");
    extern __shared__ uint32_t lds[];  // 4 quads x kGcStride
    __shared__ uint32_t red[8][8];     // per wave: packed (a, b, c) of its quad's four rows
    const int tid = threadIdx.x, ln = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), qd = wv & 3, hf = wv >> 2;
    const uint32_t j0 = blockIdx.x * kGcRows;
    if (wv < 4 && ln == 0) lds[wv * kGcStride + kGcWords] = 0xFFFFFFFFu;
    // rows j0 + 4qd + r of the wave (padding SNPs read row 0, masked)
    const uint32_t* rp[4];
    uint32_t nv[4];
    bool pad = false;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t j = j0 + 4 * qd + r;
        const uint64_t row = j < m ? (bed_idx ? bed_idx[j] : base + j) : 0;
        rp[r] = reinterpret_cast<const uint32_t*>(slabs + row * pitch);
        nv[r] = j < m ? n : 0;
        pad |= j >= m;
    }
    // the lane's 16 bytes of row r in pass hf * kHalf + ii of chunk c, from
    // a clamped address (rows are padded to 16 bytes: a 16-byte load starting
    // below src_dw stays inside the row; what lies past n is masked)
    uint4 v[kHalf][4];
    auto issue = [&](uint32_t c) {
#pragma unroll
        for (int ii = 0; ii < kHalf; ++ii) {
            const uint32_t q0 = c * cw + (hf * kHalf + ii) * 256 + 4 * ln;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                v[ii][r] = *reinterpret_cast<const uint4*>(rp[r] + (q0 < src_dw ? q0 : 0u));
        }
    };
    // per row: a = codes 01 and 11, b = codes 10 and 11, cc = code 11
    uint32_t ca[4], cb[4], cc[4];
    auto stage = [&](uint32_t c, int ii, auto mask) {
        const uint32_t k0 = (hf * kHalf + ii) * 256 + 4 * ln;  // window dword
        const uint32_t q0 = c * cw + k0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            uint32_t w[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint4 vr = v[ii][r];
                const uint32_t raw = t == 0 ? vr.x : t == 1 ? vr.y : t == 2 ? vr.z : vr.w;
                const uint32_t x = decltype(mask)::value ? ts_mask(raw, q0 + t, nv[r]) : raw;
                const uint32_t hb = (x >> 1) & 0x55555555u;  // H, at the cell's even bit
                ca[r] += __popc(x & 0x55555555u);
                cb[r] += __popc(hb);
                cc[r] += __popc(x & hb);
                w[r] = x ^ hb ^ 0xAAAAAAAAu;  // value code: (L ^ H) | (~H << 1)
            }
            uint32_t o[4];
            transpose_bytes4(w[0], w[1], w[2], w[3], o);
            reinterpret_cast<uint4*>(lds + qd * kGcStride)[gc_slot(k0 + t)] =
                make_uint4(o[0], o[1], o[2], o[3]);
        }
    };
    uint32_t* tbase = tiled + ((size_t)(j0 / kGroup) * nkb * 64 + (j0 % kGroup)) * 4;
    issue(0);
#pragma unroll 1
    for (uint32_t c = 0; c < nchunks; ++c) {
#pragma unroll
        for (int r = 0; r < 4; ++r) ca[r] = cb[r] = cc[r] = 0u;
#pragma unroll
        for (int ii = 0; ii < kHalf; ++ii) {
            const uint32_t i = hf * kHalf + ii;
            // every cell of the pass belongs to a real individual of a real SNP
            if (!pad && ((uint64_t)c * cw + (uint64_t)(i + 1) * 256) * 16 <= n)
                stage(c, ii, std::false_type{});
            else
                stage(c, ii, std::true_type{});
        }
        // wave sums, two counts a word (each below 2^14)
        const uint32_t pk[6] = {ca[0] | cb[0] << 16, ca[1] | cb[1] << 16, ca[2] | cb[2] << 16,
                                ca[3] | cb[3] << 16, cc[0] | cc[1] << 16, cc[2] | cc[3] << 16};
        uint32_t sum[6];
#pragma unroll
        for (int x = 0; x < 6; ++x) sum[x] = wave_sum_u32(pk[x]);
        if (ln == 0) {
#pragma unroll
            for (int x = 0; x < 6; ++x) red[wv][x] = sum[x];
        }
        __syncthreads();
        if (c + 1 < nchunks) issue(c + 1);  // in flight during the gather
        if (tid < kGcRows) {
            // the window's MEMBERS cells of row j: codes 00 (dosage 2), 10
            // (dosage 1), 11 (dosage 0) and 01 (missing, cells past n included)
            const int q4 = tid >> 2, r = tid & 3;
            const uint32_t j = j0 + tid;
            const uint32_t ab0 = red[q4][r], ab1 = red[q4 + 4][r];
            const uint32_t cc0 = red[q4][4 + (r >> 1)], cc1 = red[q4 + 4][4 + (r >> 1)];
            const uint32_t a = (ab0 & 0xFFFFu) + (ab1 & 0xFFFFu), b = (ab0 >> 16) + (ab1 >> 16);
            const uint32_t c3 = ((cc0 >> (16 * (r & 1))) & 0xFFFFu) + ((cc1 >> (16 * (r & 1))) & 0xFFFFu);
            const uint32_t n0 = c3, n1 = b - c3, n2 = MEMBERS - a - b + c3;
            if (j < m_pad)
                reinterpret_cast<uint4*>(part)[(size_t)c * m_pad + j] =
                    make_uint4(2 * n2 + n1, n2 + n1 + n0, 4 * n2 + n1, n1);
        }
        // gather: output dwords ql = tid + 512u of the chunk, each lane octet
        // one K-block, lane x + 4h of the octet its dword 2(x/2) + (x%2) of
        // slot half h (position dword dq = 4(x/2) + 2h + x%2).  Member e = 4k
        // + f of an output dword goes to bits 2e (the rows' own cell order):
        // its four-row word, rotated so that its field lands at bits 2f of
        // each byte, is inserted into acc[k] before the byte transpose.  A
        // 4 x 4 transpose across the lane quad then gives lane x the four
        // dwords of row 4Q + x in its slot, stored as one 16-byte piece (the
        // quad's four rows: 64 contiguous bytes).
        const uint32_t t8 = tid & 7u, xl = t8 & 3u, hh = t8 >> 2;
#pragma unroll 1
        for (int u = 0; u < kHalf; ++u) {
            const uint32_t ql = tid + kGcThreads * u;
            const uint32_t q = c * cw + (ql & ~7u) + 4 * (xl >> 1) + 2 * hh + (xl & 1);
            if (q >= nkb * 8) break;
            const uint4* lp = reinterpret_cast<const uint4*>(local + (size_t)q * 16);
            const uint4 e0 = lp[0], e1 = lp[1];
            const uint32_t ev[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
            uint32_t widx[16], rot[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {  // gc_code: LDS word * 4 + field
                const uint32_t code = (ev[e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
                widx[e] = code >> 2;
                rot[e] = (2 * (code & 3) - 2 * (e & 3)) & 31u;
            }
            // dosage-value tiles: row dword 2t + h of K-block kb -> lane
            // s + 32(t%2), dword 2(t/2) + h
            uint4* ob = reinterpret_cast<uint4*>(tbase) + (size_t)(q / 8) * 64 + 32 * hh + xl;
#pragma unroll
            for (int Q = 0; Q < 4; ++Q) {
                const uint32_t* quad = lds + Q * kGcStride;
                uint32_t acc[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const uint32_t y =
                        __builtin_amdgcn_alignbit(quad[widx[e]], quad[widx[e]], rot[e]);
                    const uint32_t mk = 0x03030303u << (2 * (e & 3));
                    acc[e >> 2] = (acc[e >> 2] & ~mk) | (y & mk);
                }
                uint32_t o[4];
                transpose_bytes4(acc[0], acc[1], acc[2], acc[3], o);  // o[r] = row 4Q + r
                // lane x keeps o[r] where r % 2 = x % 2, else takes the partner's
                // o[r ^ 1] (lanes x ^ 1), then the same at distance 2
                uint32_t b[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t pr = (uint32_t)__builtin_amdgcn_mov_dpp((int)o[r ^ 1], 0xB1, 0xF, 0xF, true);
                    b[r] = (uint32_t)(r & 1) == (xl & 1) ? o[r] : pr;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t pr = (uint32_t)__builtin_amdgcn_mov_dpp((int)b[r ^ 2], 0x4E, 0xF, 0xF, true);
                    o[r] = (uint32_t)(r >> 1) == (xl >> 1) ? b[r] : pr;
                }
                ob[4 * Q] = j0 + 4 * Q + xl < m ? make_uint4(o[0], o[1], o[2], o[3])
                                                 : make_uint4(~0u, ~0u, ~0u, ~0u);
            }
        }
        __syncthreads();  // the quads are restaged for the next chunk
    }
}

// ------------------------------------------------------------------------
// Kernel 2: per-SNP constants
// ------------------------------------------------------------------------
// int8 form: cs = (C, S), rs = 1/sqrt(C*Q), Q = C*SS - S^2 (exact integers;
// C*Q = count^2 * centered_sum_sq).  f32 form: the 4-entry LUT of
// normalize_col_f32_with_stats (normalize.rs:115-154) indexed by the PLINK
// code: [(2-mean)inv, 0, (1-mean)inv, (0-mean)inv] in f32.
// `part` holds `nparts` slots of m_pad x 4 integer stats (K-block chunks);
// their sum is written to `stats` when it is not `part` itself.
// `maf` (nullable): the MAF of normalize.rs:71-81 from the same integers,
// rounded through f32 when `maf32` (the FP32 form, normalize.rs:153).
__global__ void k_snp_info(const uint32_t* __restrict__ part, uint32_t nparts,
                           uint32_t* __restrict__ stats, uint32_t m_pad, uint32_t n,
                           int2* __restrict__ cs, double* __restrict__ rs,
                           float4* __restrict__ lut, float2* __restrict__ musc,
                           double* __restrict__ maf, int maf32) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m_pad) return;
    uint4 a = reinterpret_cast<const uint4*>(part)[j];
    for (uint32_t c = 1; c < nparts; ++c) {
        const uint4 b = reinterpret_cast<const uint4*>(part)[(size_t)c * m_pad + j];
        a.x += b.x;
        a.y += b.y;
        a.z += b.z;
        a.w += b.w;
    }
    if (stats != part) reinterpret_cast<uint4*>(stats)[j] = a;
    const uint32_t S = a.x, C = a.y, SS = a.z;
    cs[j] = make_int2((int)C, (int)S);
    const int64_t Q = (int64_t)C * (int64_t)SS - (int64_t)S * (int64_t)S;
    const int64_t V = (int64_t)C * Q;
    rs[j] = V > 0 ? 1.0 / sqrt((double)V) : 0.0;
    float a32, i32;
    const double avg = ref_mean_inv_f32(S, C, SS, n, a32, i32);
    // (0 - mean) * inv: written so that no backend fold of 0 - x into -x can
    // turn the reference's +0 (mean 0) into -0; otherwise negation is exact.
    const float w0 = a32 == 0.0f ? 0.0f : -(a32 * i32);
    lut[j] = make_float4((2.0f - a32) * i32, 0.0f, (1.0f - a32) * i32, w0);
    if (musc) musc[j] = make_float2(a32, i32);
    if (maf) {
        double freq = avg / 2.0;
        freq = freq < 0.0 ? 0.0 : (freq > 1.0 ? 1.0 : freq);
        const double v = freq < 1.0 - freq ? freq : 1.0 - freq;
        maf[j] = maf32 ? (double)(float)v : v;
    }
}

// A_eff (m_pad rows of Kp = aeff_stride(K) doubles, annotations contiguous)
// from the uploaded annotation `src` (ones when null): NaN -> 0 (read_annot,
// parse.rs:224), times pq_j when given (compute.rs:1595-1616); padding rows
// and columns are zero.  The launch covers rows j_first .. (a multiple of
// 64) of which `src` holds, column-major with column stride `stride`, the
// rows from j_first on.  A 64-SNP x 16-annotation block goes through LDS,
// so that both the reads (along SNPs) and the writes (along annotations)
// are coalesced.
__global__ __launch_bounds__(256) void k_aeff(double* __restrict__ out, const double* __restrict__ src,
                                              uint64_t stride, uint32_t j_first,
                                              const double* __restrict__ pq, uint32_t m, uint32_t K,
                                              uint32_t Kp) {
    __shared__ double blk[16][65];
    const uint32_t j0 = j_first + blockIdx.x * 64, k0 = blockIdx.y * 16, t = threadIdx.x;
    if (Kp == 1) {  // one annotation: 64 SNPs, no transposition
        if (t < 64) {
            const uint32_t j = j0 + t;
            double v = 0.0;
            if (j < m) {
                v = src ? src[j - j_first] : 1.0;
                if (v != v) v = 0.0;
                if (pq) v *= pq[j];
            }
            out[j] = v;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t kk = t / 64 + 4 * i, jj = t % 64, j = j0 + jj, k = k0 + kk;
        double v = 0.0;
        if (j < m && k < K) {
            v = src ? src[(size_t)k * stride + (j - j_first)] : 1.0;
            if (v != v) v = 0.0;
            if (pq) v *= pq[j];
        }
        blk[kk][jj] = v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t kk = t % 16, jj = t / 16 + 16 * i;
        out[(size_t)(j0 + jj) * Kp + k0 + kk] = blk[kk][jj];
    }
}

__device__ __forceinline__ void finish_2w(const BandArgs& a, EpiShared2& sh, int tile, uint32_t I,
                                          uint32_t J, int w, int lane, const v16i (&acc)[4][2]);

// r2u of one pair from its exact products G.G, G.M, M.G, M.M (integers in
// doubles), the (C, S) of its row and column SNPs and their 1/sqrt(C Q):
//   T  = C_r (C_c GG - S_c GM) - S_r (C_c MG - S_c MM)  = C_r C_c * cov
//   r  = T / sqrt(C_r Q_r C_c Q_c),  r2u = r^2 (1 + 1/D) - 1/D
// Every operation is spelled out (no contraction left to the compiler), so
// both kernels round identically; u and v are exact (integers below 2^53).
__device__ __forceinline__ double exact_r2u(const BandArgs& a, double gg, double gm, double mg,
                                            double mm, int2 cr, int2 cc, double rsr, double rsc) {
    const double u = __dsub_rn(__dmul_rn((double)cc.x, gg), __dmul_rn((double)cc.y, gm));
    const double v = __dsub_rn(__dmul_rn((double)cc.x, mg), __dmul_rn((double)cc.y, mm));
    const double T = __fma_rn((double)cr.x, u, -__dmul_rn((double)cr.y, v));
    const double rr = __dmul_rn(__dmul_rn(T, rsr), rsc);
    return __fma_rn(__dmul_rn(rr, rr), a.r2u_a, a.r2u_b);
}

// int8 form, two waves per 64x64 tile: wave w owns rows 32w .. 32w+31 (SNP
// group 2I + w) against both column groups; 4 products x 2 blocks of
// 32x32 int32 accumulators (128 registers), so two waves fit per SIMD and
// one wave's code expansion overlaps the other's MFMAs.
__device__ __forceinline__ void band_i8_tile(const BandArgs& a, EpiShared2& sh, int tile) {
    const int2 tl = a.tiles[tile];
    const uint32_t I = (uint32_t)tl.x, J = (uint32_t)tl.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t gstride = (size_t)a.nkb * 64;
    const u32x4* pa = a.tiled + (size_t)(2 * I + w) * gstride + lane;
    const u32x4* pb0 = a.tiled + (size_t)(2 * J) * gstride + lane;
    const u32x4* pb1 = pb0 + gstride;

    v16i acc[4][2];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[p][bj][r] = 0;

    u32x4 ca = pa[0], cb0 = pb0[0], cb1 = pb1[0];
    for (uint32_t kb = 0; kb < a.nkb; ++kb) {
        const size_t nk = (size_t)(kb + 1 < a.nkb ? kb + 1 : kb) * 64;
        const u32x4 na = pa[nk], nb0 = pb0[nk], nb1 = pb1[nk];
        // keep the next K-block's loads here, a full block of MFMA work ahead
        // of their use (the scheduler would otherwise sink them to the use)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            v4i ga, ma, gb[2], mb[2];
            expand_codes(ca[t], ga, ma);
            expand_codes(cb0[t], gb[0], mb[0]);
            expand_codes(cb1[t], gb[1], mb[1]);
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                acc[0][bj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ga, gb[bj], acc[0][bj], 0, 0, 0);
                acc[1][bj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ga, mb[bj], acc[1][bj], 0, 0, 0);
                acc[2][bj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ma, gb[bj], acc[2][bj], 0, 0, 0);
                acc[3][bj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ma, mb[bj], acc[3][bj], 0, 0, 0);
            }
        }
        ca = na;
        cb0 = nb0;
        cb1 = nb1;
    }

    finish_2w(a, sh, tile, I, J, w, lane, acc);
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_band_i8(BandArgs a) {
    __shared__ EpiShared2 sh;
    const int tile = tile_of_block(a);
    if (tile >= 0) band_i8_tile(a, sh, tile);
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_band_i8_batch(BandBatch b) {
    __shared__ EpiShared2 sh;
    uint32_t r;
    int tile;
    if (batch_tile(b, r, tile)) band_i8_tile(b.args[r], sh, tile);
}

// FP64 finish of wave w of a two-wave int8 tile from its exact products
// G.G, G.M, M.G, M.M, then the per-tile partial sums.
__device__ __forceinline__ void finish_2w(const BandArgs& a, EpiShared2& sh, int tile, uint32_t I,
                                          uint32_t J, int w, int lane, const v16i (&acc)[4][2]) {
    if (w == 0) {
        sh.csr[lane] = a.cs[(size_t)I * kTile + lane];
        sh.rsr[lane] = a.rs[(size_t)I * kTile + lane];
        sh.lor[lane] = a.lo[(size_t)I * kTile + lane];
    } else {
        sh.csc[lane] = a.cs[(size_t)J * kTile + lane];
        sh.rsc[lane] = a.rs[(size_t)J * kTile + lane];
    }
    __syncthreads();
    double r2u[2][16];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 32 * w + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = 32 * bj + (lane & 31);
            const int2 cr = sh.csr[row], cc = sh.csc[col];
            const double val = exact_r2u(a, (double)acc[0][bj][r], (double)acc[1][bj][r],
                                         (double)acc[2][bj][r], (double)acc[3][bj][r], cr, cc,
                                         sh.rsr[row], sh.rsc[col]);
            const uint32_t jj = I * kTile + row, ii = J * kTile + col;
            r2u[bj][r] = (ii < jj && jj < a.m && ii >= sh.lor[row]) ? val : 0.0;
        }
    band_contract2(a, sh, tile, I, J, w, r2u);
}



// Sub-blocks of a 64x64 tile: bit 2*bi + bj stands for rows 32bi.. x
// columns 32bj.. .  VcodeBlock holds one K-block (two K-steps of 64
// individuals) of the raw value codes of the groups MASK needs; VcodeOps the
// e2m1 operands X, Q of one K-step, whose MFMAs are X.X, X.Q, Q.X, Q.Q.
template <int MASK>
struct VcodeBlock {
    static constexpr bool kA0 = (MASK & 0x3) != 0, kA1 = (MASK & 0xC) != 0;
    static constexpr bool kB0 = (MASK & 0x5) != 0, kB1 = (MASK & 0xA) != 0;
    static constexpr int kGroups = (int)kA0 + (int)kA1 + (int)kB0 + (int)kB1;
    static constexpr int kMfma = 4 * (((MASK >> 0) & 1) + ((MASK >> 1) & 1) + ((MASK >> 2) & 1) +
                                      ((MASK >> 3) & 1));
    u32x4 a0, a1, b0, b1;
    // Block k; with GUARD, zeros past the last block (value code 0 in every
    // cell adds nothing to any product), so the K loop can run in whole
    // unrolled steps.
    template <bool GUARD>
    __device__ __forceinline__ void load(const u32x4* pa0, const u32x4* pa1, const u32x4* pb0,
                                         const u32x4* pb1, uint32_t k, uint32_t nkb, int lane) {
        if (!GUARD || k < nkb) {
            const size_t off = (size_t)k * 64;
            if (kA0) a0 = (pa0 + off)[lane];
            if (kA1) a1 = (pa1 + off)[lane];
            if (kB0) b0 = (pb0 + off)[lane];
            if (kB1) b1 = (pb1 + off)[lane];
        } else {
            a0 = a1 = b0 = b1 = u32x4{0u, 0u, 0u, 0u};
        }
    }
};

template <int MASK>
struct VcodeOps {
    using B = VcodeBlock<MASK>;
    v8i xa[2], qa[2], xb[2], qb[2];
    __device__ __forceinline__ void expand(const B& blk, int u) {
        if (B::kA0) vcode_operands(blk.a0[2 * u], blk.a0[2 * u + 1], xa[0], qa[0]);
        if (B::kA1) vcode_operands(blk.a1[2 * u], blk.a1[2 * u + 1], xa[1], qa[1]);
        if (B::kB0) vcode_operands(blk.b0[2 * u], blk.b0[2 * u + 1], xb[0], qb[0]);
        if (B::kB1) vcode_operands(blk.b1[2 * u], blk.b1[2 * u + 1], xb[1], qb[1]);
    }
    __device__ __forceinline__ void mma(v16f (&acc)[4][2][2]) const {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                if ((MASK >> (2 * bi + bj)) & 1) {
                    acc[0][bi][bj] = LDSC_MFMA_FP4(xa[bi], xb[bj], acc[0][bi][bj]);
                    acc[1][bi][bj] = LDSC_MFMA_FP4(xa[bi], qb[bj], acc[1][bi][bj]);
                    acc[2][bi][bj] = LDSC_MFMA_FP4(qa[bi], xb[bj], acc[2][bi][bj]);
                    acc[3][bi][bj] = LDSC_MFMA_FP4(qa[bi], qb[bj], acc[3][bi][bj]);
                }
    }
};

// One pipeline phase: the MFMAs of the expanded step `cur` run while the
// next step is expanded into `nxt`; the scheduler is asked to alternate one
// MFMA with a share of the expansion's VALU work so that the VALU issues in
// the shadow of the matrix pipe.
template <int MASK>
__device__ __forceinline__ void vcode_phase(VcodeOps<MASK>& nxt, const VcodeBlock<MASK>& src,
                                            int u, const VcodeOps<MASK>& cur,
                                            v16f (&acc)[4][2][2]) {
    using B = VcodeBlock<MASK>;
    nxt.expand(src, u);
    cur.mma(acc);
    constexpr int kValu = (14 * B::kGroups + B::kMfma - 1) / B::kMfma;
#pragma unroll
    for (int i = 0; i < B::kMfma; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);      // one MFMA
        __builtin_amdgcn_sched_group_barrier(0x002, kValu, 0);  // then VALU
    }
    __builtin_amdgcn_sched_barrier(0);
}

// K loop, software-pipelined at the K-step level and unrolled by three
// K-blocks: raw sets sA, sB, sC hold blocks kb, kb+1, kb+2 and operand sets
// e0, e1 alternate statically (no register copies).  Each raw set is
// reloaded right after its second K-step is expanded, 2.5 blocks of MFMA
// work before its next use; each K-step is expanded while the previous
// one's MFMAs run.  The block count is rounded up to a multiple of three
// with zero blocks (at most two, < 1 % of the work at panel scale), so there
// is no separate tail.  pa*/pb* are the wave-uniform starts of the tile's
// four SNP groups; each lane adds its own 16-byte slot.
template <int MASK, bool GUARD>
__device__ __forceinline__ void vcode_triple(VcodeBlock<MASK>& sA, VcodeBlock<MASK>& sB,
                                             VcodeBlock<MASK>& sC, VcodeOps<MASK>& e0,
                                             VcodeOps<MASK>& e1, const u32x4* pa0,
                                             const u32x4* pa1, const u32x4* pb0,
                                             const u32x4* pb1, uint32_t kb, uint32_t nkb,
                                             int lane, v16f (&acc)[4][2][2]) {
    vcode_phase(e1, sA, 1, e0, acc);
    sA.template load<GUARD>(pa0, pa1, pb0, pb1, kb + 3, nkb, lane);
    vcode_phase(e0, sB, 0, e1, acc);
    vcode_phase(e1, sB, 1, e0, acc);
    sB.template load<GUARD>(pa0, pa1, pb0, pb1, kb + 4, nkb, lane);
    vcode_phase(e0, sC, 0, e1, acc);
    vcode_phase(e1, sC, 1, e0, acc);
    sC.template load<GUARD>(pa0, pa1, pb0, pb1, kb + 5, nkb, lane);
    vcode_phase(e0, sA, 0, e1, acc);
}

template <int MASK>
__device__ __forceinline__ void vcode_kloop(const u32x4* pa0, const u32x4* pa1, const u32x4* pb0,
                                            const u32x4* pb1, uint32_t nkb, int lane,
                                            v16f (&acc)[4][2][2]) {
    VcodeBlock<MASK> sA, sB, sC;
    VcodeOps<MASK> e0, e1;
    sA.template load<true>(pa0, pa1, pb0, pb1, 0, nkb, lane);
    sB.template load<true>(pa0, pa1, pb0, pb1, 1, nkb, lane);
    sC.template load<true>(pa0, pa1, pb0, pb1, 2, nkb, lane);
    e0.expand(sA, 0);
    // e0 = step (kb, 0); sA, sB, sC = blocks kb .. kb+2
    uint32_t kb = 0;
    for (; kb + 5 < nkb; kb += 3)  // every block it loads exists
        vcode_triple<MASK, false>(sA, sB, sC, e0, e1, pa0, pa1, pb0, pb1, kb, nkb, lane, acc);
    for (; kb < nkb; kb += 3)      // the last one or two steps, zero blocks past the end
        vcode_triple<MASK, true>(sA, sB, sC, e0, e1, pa0, pa1, pb0, pb1, kb, nkb, lane, acc);
}

// e2m1 form of the exact kernel: one wave per 64x64 tile, 4 products x 4
// sub-blocks of 32x32 f32 accumulators (256 registers, one wave per SIMD),
// so every operand group is loaded and expanded once per K-step and feeds 8
// MFMAs.  Tiles hold 2-bit dosage-value codes v (3 = missing or padding);
// with q = [v = 3] the genotype is G = v - 3q and the presence M = 1 - q,
// so the four exact products follow from X.X, X.Q, Q.X, Q.Q on
// v_mfma_scale_f32_32x32x64_f8f6f4 (scale 1, f32 accumulation, exact while
// 9 N_pad <= 2^24):
//   VV = 4 acc0, VQ = 4 acc1, QV = 4 acc2, QQ = 4 acc3
//   GG = VV - 3 VQ - 3 QV + 9 QQ      GM = S_r - VQ + 3 QQ
//   MG = S_c - QV + 3 QQ              MM = C_r + C_c - N_pad + QQ
// Sub-blocks that cannot hold an included pair (all columns precede the
// smallest pair bound of the rows, or none precedes the last row) are
// skipped when the tile matches one of the frequent patterns (diagonal
// tile, band edges); skipped accumulators stay zero and the pair mask drops
// every pair they cover.
__device__ __forceinline__ void band_fp4_tile(const BandArgs& a, EpiShared& sh, int tile) {
    const int2 tl = a.tiles[tile];
    const uint32_t I = (uint32_t)tl.x, J = (uint32_t)tl.y;
    const int lane = threadIdx.x;
    const size_t gstride = (size_t)a.nkb * 64;
    const u32x4* pa0 = a.tiled + (size_t)(2 * I) * gstride;
    const u32x4* pb0 = a.tiled + (size_t)(2 * J) * gstride;

    const uint32_t r0 = I * kTile, c0 = J * kTile;
    uint32_t lmin = a.lo[r0 + lane];  // lanes 0-31: rows of half 0, 32-63: half 1
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) lmin = min(lmin, (uint32_t)__shfl_xor((int)lmin, o));
    const uint32_t lmin0 = __builtin_amdgcn_readlane(lmin, 0);
    const uint32_t lmin1 = __builtin_amdgcn_readlane(lmin, 32);
    int mask = 0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const uint32_t cmin = c0 + 32 * bj, rmax = r0 + 32 * bi + 31;
            if (cmin + 31 >= (bi ? lmin1 : lmin0) && cmin < rmax) mask |= 1 << (2 * bi + bj);
        }

    v16f acc[4][2][2];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[p][bi][bj][r] = 0.0f;
    const u32x4* pa1 = pa0 + gstride;
    const u32x4* pb1 = pb0 + gstride;
    if (mask == 0xF)
        vcode_kloop<0xF>(pa0, pa1, pb0, pb1, a.nkb, lane, acc);
    else if ((mask & ~0x2) == 0 && mask)
        vcode_kloop<0x2>(pa0, pa1, pb0, pb1, a.nkb, lane, acc);
    else if ((mask & ~0xD) == 0 && mask)
        vcode_kloop<0xD>(pa0, pa1, pb0, pb1, a.nkb, lane, acc);
    else if ((mask & ~0xB) == 0 && mask)
        vcode_kloop<0xB>(pa0, pa1, pb0, pb1, a.nkb, lane, acc);
    else if (mask)
        vcode_kloop<0xF>(pa0, pa1, pb0, pb1, a.nkb, lane, acc);

    stage_info(a, sh, I, J);
    __syncthreads();
    double r2u[2][2][16];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = row_of(bi, r, lane), col = 32 * bj + (lane & 31);
                const int2 cr = sh.csr[row], cc = sh.csc[col];
                // exact integers
                const double vv = 4.0 * (double)acc[0][bi][bj][r], vq = 4.0 * (double)acc[1][bi][bj][r];
                const double qv = 4.0 * (double)acc[2][bi][bj][r], qq = 4.0 * (double)acc[3][bi][bj][r];
                const double gg = vv - 3.0 * (vq + qv) + 9.0 * qq;  // all exact
                const double gm = (double)cr.y - vq + 3.0 * qq;
                const double mg = (double)cc.y - qv + 3.0 * qq;
                const double mm = (double)cr.x + (double)cc.x - a.n_pad + qq;
                r2u[bi][bj][r] = pair_included(a, sh, I, J, row, col)
                                     ? exact_r2u(a, gg, gm, mg, mm, cr, cc, sh.rsr[row], sh.rsc[col])
                                     : 0.0;
            }
    band_contract(a, sh, tile, I, J, r2u);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_band_fp4(BandArgs a) {
    __shared__ EpiShared sh;
    const int tile = tile_of_block(a);
    if (tile >= 0) band_fp4_tile(a, sh, tile);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_band_fp4_batch(BandBatch b) {
    __shared__ EpiShared sh;
    uint32_t r;
    int tile;
    if (batch_tile(b, r, tile)) band_fp4_tile(b.args[r], sh, tile);
}

// Standardized f32 operands come from the tile's LUTs in LDS as pairs: a
// 4-bit field of a code dword holds the codes of two individuals (fields f
// and f + 1 of one byte), and entry c = 0 .. 15 of a SNP's pair table is
// (L[c & 3], L[c >> 2]) of its LUT L = [(2-mu)inv, 0, (1-mu)inv, (0-mu)inv].
// Code major: entry c of slot x (rows 0-63, then columns 64-127) is float2
// 128 c + x.  A lane reads its own slot, so two operands cost a bit-field
// extract, a shift-or and one 8-byte LDS read, and the 32 lanes of a
// half-wave read 32 consecutive float2 for any mix of codes.
constexpr uint32_t kLutSlots = 128;
static_assert(sizeof(EpiShared2::u) >= 16 * kLutSlots * sizeof(float2), "pair tables in the epilogue LDS");
__device__ __forceinline__ float2 pair_pick(const float2* slot, uint32_t w, uint32_t sh) {
    return slot[__builtin_amdgcn_ubfe(w, sh, 4u) * kLutSlots];
}

// K loop of one wave of the two-wave FP32 kernel: its row group against the
// column groups MASK names (bit bj: columns 32bj ..), one
// v_mfma_f32_32x32x2f32 per column group and pair of individuals; la, lb0,
// lb1 are the lane's pair-table slots (its row, its column in each group).
template <int MASK>
__device__ __forceinline__ void f32_2w_kloop(const u32x4* pa, const u32x4* pb0, const u32x4* pb1,
                                             const float2* la, const float2* lb0,
                                             const float2* lb1, uint32_t nkb, v16f (&acc)[2]) {
    constexpr bool kB0 = (MASK & 1) != 0, kB1 = (MASK & 2) != 0;
    const u32x4 z = {0u, 0u, 0u, 0u};
    u32x4 ca = pa[0], cb0 = kB0 ? pb0[0] : z, cb1 = kB1 ? pb1[0] : z;
    for (uint32_t kb = 0; kb < nkb; ++kb) {
        const size_t nk = (size_t)(kb + 1 < nkb ? kb + 1 : kb) * 64;
        const u32x4 na = pa[nk], nb0 = kB0 ? pb0[nk] : z, nb1 = kB1 ? pb1[nk] : z;
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch a block ahead
        // per code dword: all its operands first, then its MFMAs, so that
        // the scheduler runs one dword's MFMAs beside the next one's table
        // reads (the same MFMA sequence as picking each operand just before
        // its MFMA).  Individual s of the dword is field s >> 2 of byte
        // s & 3, so nibble j of byte k holds individuals k + 8j and k + 8j + 4.
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            float fa[16], fb0[16], fb1[16];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int k = q & 3, j = q >> 2, s0 = k + 8 * j;
                const uint32_t shn = 8 * k + 4 * j;
                const float2 pa2 = pair_pick(la, ca[t], shn);
                fa[s0] = pa2.x;
                fa[s0 + 4] = pa2.y;
                if (kB0) {
                    const float2 pb = pair_pick(lb0, cb0[t], shn);
                    fb0[s0] = pb.x;
                    fb0[s0 + 4] = pb.y;
                }
                if (kB1) {
                    const float2 pb = pair_pick(lb1, cb1[t], shn);
                    fb1[s0] = pb.x;
                    fb1[s0 + 4] = pb.y;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                if (kB0) acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[s], fb0[s], acc[0], 0, 0, 0);
                if (kB1) acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[s], fb1[s], acc[1], 0, 0, 0);
            }
        }
        ca = na;
        cb0 = nb0;
        cb1 = nb1;
    }
}

// FP32 form, two waves per 64x64 tile: wave w owns rows 32w .. 32w+31 (SNP
// group 2I + w) against both column groups (2 blocks of 32x32 f32
// accumulators), so two waves share a SIMD and one wave's LUT reads and
// FP64 epilogue run beside the other's MFMAs.  Each wave skips a 32x32
// block of its rows that cannot hold an included pair.  Each dot product
// is the same sequence of MFMAs over the individuals as one wave per tile
// would run.
__device__ __forceinline__ void band_f32_2w_tile(const BandArgs& a, EpiShared2& sh, int tile) {
    const int2 tl = a.tiles[tile];
    const uint32_t I = (uint32_t)tl.x, J = (uint32_t)tl.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t gstride = (size_t)a.nkb * 64;
    const u32x4* pa = a.tiled + (size_t)(2 * I + w) * gstride + lane;
    const u32x4* pb0 = a.tiled + (size_t)(2 * J) * gstride + lane;
    const u32x4* pb1 = pb0 + gstride;
    const uint32_t r0 = I * kTile + 32 * w, c0 = J * kTile;
    // the tile's pair tables in LDS, in the epilogue's storage (unused until
    // the K loop ends): wave 0 stores the row SNPs' (slots 0-63), wave 1 the
    // column SNPs' (slots 64-127)
    float2* pair_s = reinterpret_cast<float2*>(&sh.u);
    {
        const float4 v = a.lut[(w ? c0 : I * kTile) + lane];
        const float L4[4] = {v.x, v.y, v.z, v.w};
        float2* d = pair_s + 64 * w + lane;
#pragma unroll
        for (int c = 0; c < 16; ++c) d[c * kLutSlots] = make_float2(L4[c & 3], L4[c >> 2]);
    }
    __syncthreads();
    const float2* la = pair_s + 32 * w + (lane & 31);
    const float2* lb0 = pair_s + 64 + (lane & 31);
    const float2* lb1 = pair_s + 96 + (lane & 31);

    uint32_t lmin = a.lo[r0 + (lane & 31)];
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) lmin = min(lmin, (uint32_t)__shfl_xor((int)lmin, o));
    lmin = __builtin_amdgcn_readlane(lmin, 0);
    int mask = 0;
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
        const uint32_t cmin = c0 + 32 * bj;
        if (cmin + 31 >= lmin && cmin < r0 + 31) mask |= 1 << bj;
    }

    v16f acc[2];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[bj][r] = 0.0f;
    if (mask == 3)
        f32_2w_kloop<3>(pa, pb0, pb1, la, lb0, lb1, a.nkb, acc);
    else if (mask == 1)
        f32_2w_kloop<1>(pa, pb0, pb1, la, lb0, lb1, a.nkb, acc);
    else if (mask == 2)
        f32_2w_kloop<2>(pa, pb0, pb1, la, lb0, lb1, a.nkb, acc);

    if (w == 0) sh.lor[lane] = a.lo[(size_t)I * kTile + lane];
    __syncthreads();
    // compute.rs:2064-2077: val (f32 GEMM output) -> f64, val^2 * n_inv^2 * r2u_a + r2u_b
    double r2u[2][16];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 32 * w + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = 32 * bj + (lane & 31);
            const double v = (double)acc[bj][r];
            const double val = __dadd_rn(__dmul_rn(__dmul_rn(v, v), a.n_inv_sq_r2u_a), a.r2u_b);
            const uint32_t jj = I * kTile + row, ii = J * kTile + col;
            r2u[bj][r] = (ii < jj && jj < a.m && ii >= sh.lor[row]) ? val : 0.0;
        }
    band_contract2(a, sh, tile, I, J, w, r2u);
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_band_f32_2w(BandArgs a) {
    __shared__ EpiShared2 sh;
    const int tile = tile_of_block(a);
    if (tile >= 0) band_f32_2w_tile(a, sh, tile);
}

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_band_f32_2w_batch(BandBatch b) {
    __shared__ EpiShared2 sh;
    uint32_t r;
    int tile;
    if (batch_tile(b, r, tile)) band_f32_2w_tile(b.args[r], sh, tile);
}

// ------------------------------------------------------------------------
// Kernel 4: deterministic per-SNP reduction
// ------------------------------------------------------------------------
// SNPs j_begin .. j_begin + j_count - 1 and annotations k_first ..
// k_first + k_count - 1; the result of (j, k) goes to
// l2[j_begin K + k j_count + (j - j_begin)]: column-major over the whole
// matrix when the range is every SNP, else a block of its own (the rows of
// a segment, column-major within it).
__global__ void k_reduce(const double* __restrict__ rowp, const double* __restrict__ colp,
                         const double* __restrict__ aeff, const int* __restrict__ j0,
                         const int* __restrict__ tstart, uint32_t nb, uint32_t j_begin,
                         uint32_t j_count, uint32_t K, uint32_t k_first, uint32_t k_count,
                         double* __restrict__ l2) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)j_count * k_count) return;
    const uint32_t k = k_first + (uint32_t)(gid / j_count), j = j_begin + (uint32_t)(gid % j_count);
    const uint32_t I = j / kTile, r = j % kTile;
    double s = aeff[(size_t)j * aeff_stride(K) + k];  // self pair, r2u(j, j) = 1 (compute.rs:1658)
    // row tiles J = j0[I] .. I, then column tiles I2 = I .. while j0[I2] <= I;
    // loads four at a time, sums strictly in tile order
    const double* pr = rowp + ((size_t)tstart[I] * K + k) * kTile + r;
    const int nr = (int)I - j0[I] + 1;
    const size_t step = (size_t)K * kTile;
    int t = 0;
    for (; t + 4 <= nr; t += 4) {
        const double a0 = pr[t * step], a1 = pr[(t + 1) * step], a2 = pr[(t + 2) * step],
                     a3 = pr[(t + 3) * step];
        s += a0;
        s += a1;
        s += a2;
        s += a3;
    }
    for (; t < nr; ++t) s += pr[t * step];
    uint32_t I2 = I;
    for (; I2 + 4 <= nb && (uint32_t)j0[I2 + 3] <= I; I2 += 4) {  // j0 is nondecreasing
        double a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            a[u] = colp[(((size_t)tstart[I2 + u] + (I - j0[I2 + u])) * K + k) * kTile + r];
#pragma unroll
        for (int u = 0; u < 4; ++u) s += a[u];
    }
    for (; I2 < nb && (uint32_t)j0[I2] <= I; ++I2)
        s += colp[(((size_t)tstart[I2] + (I - j0[I2])) * K + k) * kTile + r];
    l2[(size_t)j_begin * K + (size_t)k * j_count + (j - j_begin)] = s;
}

// ------------------------------------------------------------------------
// Decode entry point kernels
// ------------------------------------------------------------------------
__global__ void k_decode(const uint8_t* __restrict__ slabs, uint64_t bps,
                         const int64_t* __restrict__ keep, uint32_t n,
                         const uint64_t* __restrict__ bed_idx, uint32_t m,
                         const float4* __restrict__ lut, int8_t* __restrict__ dosage,
                         uint32_t* __restrict__ miss, float* __restrict__ stdv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = blockIdx.y;
    if (j >= m) return;
    const uint8_t* slab = slabs + bed_idx[j] * bps;
    uint32_t code = 1;
    if (i < n) {
        const uint64_t row = keep ? (uint64_t)keep[i] : i;
        code = (slab[row >> 2] >> (2 * (row & 3))) & 3u;
        if (dosage) dosage[(size_t)j * n + i] = (int8_t)(code == 0 ? 2 : code == 1 ? -1 : code == 2 ? 1 : 0);
        if (stdv) stdv[(size_t)j * n + i] = pick_code(lut[j], code);
    }
    if (miss) {
        const uint64_t ball = __ballot(i < n && code == 1);
        const uint32_t words = (n + 31) / 32;
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t w0 = (i - lane) / 32;  // first 32-bit word of this wave
        if (lane == 0 && w0 < words) miss[(size_t)j * words + w0] = (uint32_t)ball;
        if (lane == 32 && w0 + 1 < words) miss[(size_t)j * words + w0 + 1] = (uint32_t)(ball >> 32);
    }
}

// Per-SNP stats for an arbitrary SNP list without tiling (decode entry).
__global__ __launch_bounds__(256) void k_stats_only(const uint8_t* __restrict__ slabs, uint64_t bps,
                                                    const int64_t* __restrict__ keep, uint32_t n,
                                                    const uint64_t* __restrict__ bed_idx,
                                                    uint32_t* __restrict__ stats) {
    __shared__ uint32_t red[4][256];
    const uint32_t j = blockIdx.x;
    const uint8_t* slab = slabs + bed_idx[j] * bps;
    uint32_t c2 = 0, c1 = 0, c0 = 0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t row = keep ? (uint64_t)keep[i] : i;
        const uint32_t code = (slab[row >> 2] >> (2 * (row & 3))) & 3u;
        c2 += code == 0;
        c1 += code == 2;
        c0 += code == 3;
    }
    red[0][threadIdx.x] = 2 * c2 + c1;
    red[1][threadIdx.x] = c2 + c1 + c0;
    red[2][threadIdx.x] = 4 * c2 + c1;
    red[3][threadIdx.x] = c1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) stats[4 * j + threadIdx.x] = red[threadIdx.x][0];
}

// ------------------------------------------------------------------------
// Host orchestration
// ------------------------------------------------------------------------
static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

uint64_t band_tile_count(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                         std::vector<uint64_t>& lo) {
    if (m == 0) return 0;
    pair_bounds(snps, m, cfg, lo, nullptr);
    uint64_t tiles = 0;
    for (uint64_t I = 0; I * kTile < m; ++I) tiles += I - lo[I * kTile] / kTile + 1;
    return tiles;
}

size_t l2_workspace_bytes(const ldsc_panel* panel, uint64_t m, const ldsc_l2_config* cfg,
                          uint64_t K, uint64_t tiles, bool annotated) {
    if (m == 0) return 0;
    const uint64_t m_pad = round_up(m, kTile);
    const uint64_t n_pad = round_up(panel->n, kKBlock);
    const uint64_t nb = m_pad / kTile;
    size_t bytes = m_pad * n_pad / 4;                            // tiled SNP slice
    bytes += m_pad * 16 * (1 + tile_parts(panel->n));            // stats + chunk slots
    bytes += m * 8 + m_pad * (4 + 48 + 8 + 8);                   // idx, lo, info, kappa, maf
    bytes += m_pad * aeff_stride((uint32_t)K) * 8 + m * K * 8;   // A_eff, l2
    if (annotated) bytes += m * K * 8;                           // uploaded annotation
    if (cfg && !std::isnan(cfg->pq_exp)) bytes += m * 8;         // pq
    bytes += tiles * (kTile * K * 8 * 2 + 8) + (2 * nb + 1) * 4; // partials, tile list, plan
    const bool gather = cfg && (cfg->sketch_d || panel->d_keep32);
    if (gather) bytes += m * round_up((panel->n + 3) / 4, 4);    // permuted rows
    if (cfg && cfg->sketch_d) {
        const uint64_t nblocks = round_up(cfg->sketch_d, 32) / 32;
        const uint64_t shares = m_pad * 8 * sketch_share_slots(nblocks);
        bytes += m_pad * nblocks * 32 * 4;                       // sketch tiles
        bytes += shares + std::max<uint64_t>(m_pad * 4, shares + 3 * m_pad * 8 + 16);
    }
    // DevBuf rounds each of its ~20 buffers up to at least 256 bytes
    return bytes + 20 * 256;
}

// MAF of one SNP from exact integer stats (normalize.rs:71-81).
static double maf_from_stats(uint32_t S, uint32_t C) {
    const double avg = C > 0 ? (double)S / (double)C : 0.0;
    double freq = avg / 2.0;
    freq = freq < 0.0 ? 0.0 : (freq > 1.0 ? 1.0 : freq);
    return freq < 1.0 - freq ? freq : 1.0 - freq;
}

int gather_and_tile(const ldsc_panel* panel, const uint64_t* d_bed_idx, uint64_t base,
                    uint64_t m, const int32_t* d_map, uint64_t m_pad, bool vcode, bool tile,
                    Workspace* ws, uint32_t* parts, char* err, uint64_t err_len) {
    const uint64_t n = panel->n;
    const uint64_t n_pad = round_up(n, kKBlock);
    const uint32_t nkb = (uint32_t)(n_pad / kKBlock);
    hipStream_t st = ws->stream;
    const uint32_t nparts = tile_parts(n);
    HIPC(ws->stats.ensure((size_t)m_pad * 16 * (1 + nparts)));
    *parts = nparts;
    const uint8_t* src = panel->d_slabs;
    uint64_t pitch = panel->stride;
    const uint64_t* rows = d_bed_idx;
    const uint32_t src_dw = (uint32_t)((panel->bps + 3) / 4);
    if (d_map && (size_t)src_dw * 16 <= 160 * 1024) {
        // gather four rows at a time through byte-interleaved LDS, into the
        // tiles directly when tiles are wanted; stats go to slot 0
        const uint32_t n16 = (uint32_t)((n + 15) / 16);
        const int form = !tile ? kGatherRows : vcode ? kGatherValueTiles : kGatherCodeTiles;
        uint32_t n_out = n16;
        uint32_t* a_out;
        if (form == kGatherRows) {
            HIPC(ws->scratch.ensure((size_t)m * n16 * 4));
            a_out = ws->scratch.as<uint32_t>();
        } else {
            HIPC(ws->tiled.ensure((size_t)m_pad * n_pad / 4));
            a_out = ws->tiled.as<uint32_t>();
            n_out = nkb * 8;
        }
        const size_t lds = (size_t)src_dw * 16;
        using GatherFn = void (*)(const uint8_t*, uint64_t, uint32_t, const uint64_t*, uint64_t,
                                  uint32_t, uint32_t, const int32_t*, uint32_t, uint32_t,
                                  uint32_t*, uint32_t*);
        static const GatherFn fns[4][3] = {
            {k_gather4<4, 0>, k_gather4<4, 1>, k_gather4<4, 2>},
            {k_gather4<8, 0>, k_gather4<8, 1>, k_gather4<8, 2>},
            {k_gather4<16, 0>, k_gather4<16, 1>, k_gather4<16, 2>},
            {k_gather4<0, 0>, k_gather4<0, 1>, k_gather4<0, 2>}};
        const int pf = src_dw <= 2048 ? 0 : src_dw <= 4096 ? 1 : src_dw <= 8192 ? 2 : 3;
        const void* kfn = (const void*)fns[pf][form];
        if (lds > 64 * 1024)
            HIPC(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const uint32_t grid = (uint32_t)std::min<uint64_t>((m_pad + 3) / 4, 2048);
        const uint32_t a_m = (uint32_t)m, a_mp = (uint32_t)m_pad;
        uint32_t* a_part = ws->stats.as<uint32_t>() + (size_t)m_pad * 4;
        const uint8_t* a_sl = panel->d_slabs;
        const uint64_t a_pitch = panel->stride;
        void* kargs[] = {(void*)&a_sl, (void*)&a_pitch, (void*)&src_dw, (void*)&d_bed_idx,
                         (void*)&base, (void*)&a_m, (void*)&a_mp, (void*)&d_map, (void*)&n16,
                         (void*)&n_out, (void*)&a_out, (void*)&a_part};
        HIPC(hipLaunchKernel(kfn, dim3(grid), dim3(512), kargs, lds, st));
        *parts = 1;
        return LDSC_OK;
    } else if (d_map) {
        // gather rows in the requested individual order, then tile them
        const uint32_t n16 = (uint32_t)((n + 15) / 16);
        const uint64_t bps_out = 4ull * n16;
        const uint32_t ldsrow = (uint32_t)round_up(panel->bps, 16);
        if (ldsrow > 160 * 1024)
            return fail(err, err_len, LDSC_ERR_INVALID_ARG,
                        "individual subsets / sketching support at most 655360 FAM rows");
        HIPC(ws->scratch.ensure((size_t)m * bps_out));
        int R = (int)std::max<uint32_t>(1, (96 * 1024) / ldsrow);
        R = R >= 8 ? 8 : (R >= 4 ? 4 : (R >= 2 ? 2 : 1));
        const size_t lds = (size_t)R * ldsrow;
        const void* kfn = R == 8 ? (const void*)k_permute_rows<8>
                        : R == 4 ? (const void*)k_permute_rows<4>
                        : R == 2 ? (const void*)k_permute_rows<2> : (const void*)k_permute_rows<1>;
        if (lds > 64 * 1024)
            HIPC(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const uint32_t grid = (uint32_t)std::min<uint64_t>((m + R - 1) / R, 4096);
        const uint8_t* a0 = panel->d_slabs;
        uint8_t* a8 = ws->scratch.as<uint8_t>();
        uint64_t a1 = panel->stride;
        uint32_t a2 = (uint32_t)((panel->bps + 3) / 4), a5 = (uint32_t)m;
        uint64_t a_base = base;
        void* kargs[] = {(void*)&a0, (void*)&a1, (void*)&a2, (void*)&ldsrow, (void*)&d_bed_idx,
                         (void*)&a_base, (void*)&a5, (void*)&d_map, (void*)&n16, (void*)&a8};
        HIPC(hipLaunchKernel(kfn, dim3(grid), dim3(512), kargs, lds, st));
        src = ws->scratch.as<uint8_t>();
        pitch = bps_out;
        rows = nullptr;
        base = 0;
    }
    HIPC(ws->tiled.ensure((size_t)m_pad * n_pad / 4));
    hipLaunchKernelGGL(vcode ? k_tile_stats<true> : k_tile_stats<false>,
                       dim3((unsigned)(m_pad / kGroup), nparts), dim3(256), 0, st, src, pitch,
                       (uint32_t)n, rows, base, (uint32_t)m, nkb, (uint32_t)m_pad,
                       ws->tiled.as<uint32_t>(), ws->stats.as<uint32_t>() + (size_t)m_pad * 4);
    HIPC(hipGetLastError());
    return LDSC_OK;
}

int gather_tile_chunks(const ldsc_panel* panel, const uint64_t* d_bed_idx, uint64_t base,
                       uint64_t m, const uint16_t* d_local, uint32_t chunk, uint64_t m_pad,
                       Workspace* ws, uint32_t* parts, char* err, uint64_t err_len) {
    const uint64_t n = panel->n;
    const uint64_t n_pad = round_up(n, kKBlock);
    const uint32_t nkb = (uint32_t)(n_pad / kKBlock);
    if (chunk != kSkMembers)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG, "unsupported sketch chunk %u", chunk);
    const uint32_t nchunks = (uint32_t)((n + chunk - 1) / chunk);
    hipStream_t st = ws->stream;
    HIPC(ws->stats.ensure((size_t)m_pad * 16 * (1 + nchunks)));
    HIPC(ws->tiled.ensure((size_t)m_pad * n_pad / 4));
    *parts = nchunks;
    const size_t lds = (size_t)4 * (chunk / 4 + 4) * 4;  // four row quads
    HIPC(hipFuncSetAttribute((const void*)k_gather_chunk<kSkMembers>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_gather_chunk<kSkMembers>, dim3((unsigned)(m_pad / kGcRows)),
                       dim3(kGcThreads), lds,
                       st, panel->d_slabs, panel->stride, (uint32_t)((panel->bps + 3) / 4),
                       (uint32_t)n, d_bed_idx, base, (uint32_t)m, (uint32_t)m_pad, d_local, nkb,
                       nchunks, ws->tiled.as<uint32_t>(),
                       ws->stats.as<uint32_t>() + (size_t)m_pad * 4);
    HIPC(hipGetLastError());
    return LDSC_OK;
}

// MAF of the call's SNPs on the host from the exact integer stats in
// pinned memory (the sketch form, whose rounding depends on chunk
// contiguity).  f32 exact path: (f32)maf (normalize.rs:153).  Sketch: f64
// on the fused path of a BED-contiguous chunk (compute.rs:639), (f32)maf
// otherwise (compute.rs:1509-1511).
static void compute_maf_host(const L2Call& c) {
    const uint64_t m = c.m;
    const ldsc_snp* snps = c.snps;
    const uint32_t* h_stats =
        reinterpret_cast<const uint32_t*>(c.ws->pinned_meta.as<char>() + c.off_st);
    std::vector<double>& maf = c.ws->host.maf;
    maf.resize(m);
    std::vector<uint8_t>& round32 = c.ws->host.round32;
    round32.assign(m, c.use_f32 && !c.d ? 1 : 0);
    if (c.d) {
        uint64_t run0 = 0;
        while (run0 < m) {
            uint64_t run1 = m;
            if (c.cfg->per_chromosome) {
                run1 = run0 + 1;
                while (run1 < m && snps[run1].chr == snps[run0].chr) ++run1;
            }
            for (uint64_t s0 = run0; s0 < run1; s0 += c.cfg->chunk_size) {
                const uint64_t e = std::min<uint64_t>(s0 + c.cfg->chunk_size, run1);
                const bool contig = snps[e - 1].bed_idx == snps[s0].bed_idx + (e - s0) - 1;
                for (uint64_t j = s0; j < e; ++j) round32[j] = contig ? 0 : 1;
            }
            run0 = run1;
        }
    }
    for (uint64_t j = 0; j < m; ++j) {
        const double v = maf_from_stats(h_stats[4 * j], h_stats[4 * j + 1]);
        maf[j] = round32[j] ? (double)(float)v : v;
    }
}

int l2_begin(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
             const double* annot, uint64_t K, double* l2_out, double* maf_out, ldsc_l2_perf* perf,
             ldsc_progress_cb cb, void* cb_ctx, Workspace* ws, L2Call* call, char* err,
             uint64_t err_len, bool segment_annot) {
    const double t_start = now_s();
    L2Call& c = *call;
    c = L2Call{};
    c.done = true;
    int rc = check_config(cfg, err, err_len);
    if (rc) return rc;
    if (!panel) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "panel is NULL");
    if (K < 1) return fail(err, err_len, LDSC_ERR_ANNOT, "annotation column count must be >= 1");
    if (!annot && K != 1)
        return fail(err, err_len, LDSC_ERR_ANNOT, "K must be 1 without an annotation matrix");
    const uint64_t n = panel->n;
    const uint64_t d = cfg->sketch_d;
    if (d && !(d >= 3 && d < n))
        return fail(err, err_len, LDSC_ERR_SKETCH_RANGE,
                    "--sketch %llu is out of bounds (must be 3 <= d < n_indiv=%llu)",
                    (unsigned long long)d, (unsigned long long)n);
    if (m == 0) {
        if (perf) {
            std::memset(perf, 0, sizeof(*perf));
            perf->n_indiv = n;
        }
        return LDSC_OK;
    }
    if (!snps || !l2_out) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL SNP or output array");
    if (m >= (1ull << 31)) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "too many SNPs");
    // rows base .. base+m-1 of the panel need no index upload; the check
    // runs in parallel pieces and reports the first SNP out of range
    const uint64_t base = snps[0].bed_idx;
    bool contiguous = true;
    {
        const int T = task_count(m);
        std::vector<uint64_t> bad(T, UINT64_MAX);
        std::vector<uint8_t> contig(T, 1);
        parallel_tasks(T, [&](int t) {
            const uint64_t a = m * t / T, b = m * (t + 1) / T;
            bool ct = true;
            for (uint64_t j = a; j < b; ++j) {
                const uint64_t bi = snps[j].bed_idx;
                if (bi >= panel->m_bim) {
                    bad[t] = j;
                    break;
                }
                ct = ct && bi == base + j;
            }
            contig[t] = ct ? 1 : 0;
        });
        for (int t = 0; t < T; ++t) {
            if (bad[t] != UINT64_MAX)
                return fail(err, err_len, LDSC_ERR_INVALID_ARG,
                            "SNP %llu has bed_idx %llu beyond the panel's %llu SNPs",
                            (unsigned long long)bad[t], (unsigned long long)snps[bad[t]].bed_idx,
                            (unsigned long long)panel->m_bim);
            contiguous = contiguous && contig[t];
        }
    }
    if (d && d > 1000000)
        return fail(err, err_len, LDSC_ERR_SKETCH_RANGE, "--sketch %llu too large",
                    (unsigned long long)d);

    const uint32_t m_pad = (uint32_t)round_up(m, kTile);
    HIPC(hipSetDevice(panel->device));
    hipStream_t st = ws->stream;
    c.panel = panel;
    c.snps = snps;
    c.m = m;
    c.cfg = cfg;
    c.K = K;
    c.l2_out = l2_out;
    c.maf_out = maf_out;
    c.perf = perf;
    c.cb = cb;
    c.cb_ctx = cb_ctx;
    c.ws = ws;
    c.t_start = t_start;
    c.m_pad = m_pad;
    c.d = d;
    c.bed_contiguous = contiguous;

    SketchPlan* sp = nullptr;
    if (d) {
        rc = sketch_plan(panel, d, &sp, err, err_len);
        if (rc) return rc;
    }

    // 1. gather + tile + stats (+ CountSketch), queued before the host work
    //    below so the device tiles while the host plans the band.
    uint64_t* h_bidx = nullptr;
    if (!contiguous) {
        HIPC(ws->pinned.ensure((size_t)m * 8));
        h_bidx = ws->pinned.as<uint64_t>();
        const int T = task_count(m);
        parallel_tasks(T, [&](int t) {
            const uint64_t a = m * t / T, b = m * (t + 1) / T;
            for (uint64_t j = a; j < b; ++j) h_bidx[j] = snps[j].bed_idx;
        });
        HIPC(ws->bed_idx.ensure((size_t)m * 8));
    }
    HIPC(ws->info.ensure((size_t)m_pad * 48));
    if (d) {
        HIPC(ws->sk_cols.ensure((size_t)m_pad * sp->d_pad * 4));
        HIPC(ws->maf.ensure((size_t)m_pad * 8));  // kappa (maf-aware sketch)
    }
    int2* d_cs = ws->info.as<int2>();
    double* d_rs = reinterpret_cast<double*>(ws->info.as<char>() + (size_t)m_pad * 8);
    float4* d_lut = reinterpret_cast<float4*>(ws->info.as<char>() + (size_t)m_pad * 16);
    float2* d_musc = reinterpret_cast<float2*>(ws->info.as<char>() + (size_t)m_pad * 32);
    if (!contiguous)
        HIPC(hipMemcpyAsync(ws->bed_idx.p, h_bidx, m * 8, hipMemcpyHostToDevice, st));
    HIPC(hipEventRecord(ws->ev[0], st));
    const bool use_f32 = cfg->precision == LDSC_PREC_FP32;
    // exact form: e2m1 MFMA while f32 holds the products exactly, int8 beyond
    const bool use_fp4 = !d && cfg->precision == LDSC_PREC_EXACT && n <= kFp4MaxIndiv;
    c.use_f32 = use_f32;
    // the sketch's d-length contraction runs on exact int8 operands when its
    // int32 sums stay exact (falling back to f32 values if a value leaves
    // the int8 range, l2_band)
    const bool sk8 = d && sp->d_pad <= kSketchI8MaxDpad;
    c.form = d ? (sk8 ? kBandSketch8 : kBandSketch)
               : use_f32 ? kBandF32 : use_fp4 ? kBandFp4 : kBandI8;
    uint32_t parts = 1;
    // chunk-local bucket order in the integer form with at most 8 bucket
    // blocks (d <= 256): one pass from the .bed rows to the bucket sums, no
    // tiled copy (the range fallback of l2_band tiles the rows then); up to
    // 7 blocks in natural order (k_sketch_dense, when the plan built its
    // codes), else in bucket order (k_sketch_fused).  With more blocks a wave
    // owns several and the pass measured slower than the tiled gather plus
    // k_sketch.
    const bool fused = sk8 && sp->d_local && sp->nblocks <= kSfFusedMaxBlocks;
    c.sk_fused = fused;
    if (fused) {
        // sized for the fallback's chunk slots too, so that it never moves
        HIPC(ws->stats.ensure((size_t)m_pad * 16 * (1 + sp->nchunks)));
        HIPC(ws->sk_norm.ensure(sketch_norm_bytes(sp, m_pad)));
        HIPC(launch_sketch_fused(sp, panel, contiguous ? nullptr : ws->bed_idx.as<uint64_t>(),
                                 contiguous ? base : 0, (uint32_t)m, m_pad,
                                 ws->stats.as<uint32_t>() + (size_t)m_pad * 4,
                                 ws->sk_norm.as<double>(), ws->sk_cols.p, st));
        parts = 1;
    } else if (d && sp->d_local)
        rc = gather_tile_chunks(panel, contiguous ? nullptr : ws->bed_idx.as<uint64_t>(),
                                contiguous ? base : 0, m, sp->d_local, (uint32_t)sp->chunk, m_pad,
                                ws, &parts, err, err_len);
    else
        rc = gather_and_tile(panel, contiguous ? nullptr : ws->bed_idx.as<uint64_t>(),
                             contiguous ? base : 0, m, d ? sp->d_fam : panel->d_keep32, m_pad,
                             use_fp4 || d, /*tile=*/true, ws, &parts, err, err_len);
    if (rc) return rc;
    // MAF on the device for the exact and FP32 forms; the sketch form's
    // rounding depends on chunk contiguity and is finished on the host
    const bool dev_maf = !d;
    if (dev_maf) HIPC(ws->mafd.ensure((size_t)m_pad * 8));
    hipLaunchKernelGGL(k_snp_info, dim3((m_pad + 255) / 256), dim3(256), 0, st,
                       ws->stats.as<uint32_t>() + (size_t)m_pad * 4, parts,
                       ws->stats.as<uint32_t>(), m_pad, (uint32_t)n, d_cs, d_rs, d_lut, d_musc,
                       dev_maf ? ws->mafd.as<double>() : nullptr, use_f32 ? 1 : 0);
    HIPC(hipGetLastError());
    // the annotation goes up on its own stream, from a thread of its own,
    // while the device tiles and this thread plans the band
    std::thread upload;
    hipError_t upload_err = hipSuccess;
    struct Joiner {
        std::thread& t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } upload_joiner{upload};
    // A large annotation goes up in row segments: the first here, the others
    // in l2_band, each while the band of the segment before it runs.
    uint32_t nseg = 1;
    if (segment_annot && annot) {
        const uint64_t bytes = m * K * 8;
        nseg = bytes >= (128ull << 20) ? 4u : bytes >= (32ull << 20) ? 2u : 1u;
        if (m < (uint64_t)nseg * 1024) nseg = 1;
    }
    // The first segment's rows must be on the device before the band starts
    // and the last segment's results come back after it ends, so those two
    // are half the size of the others (row weights 1, 2, ..., 2, 1).
    uint32_t seg_r[L2Call::kMaxSeg + 1];
    for (uint32_t sg = 0; sg <= nseg; ++sg)
        seg_r[sg] = sg == nseg ? m_pad
                    : sg == 0  ? 0u
                               : (uint32_t)round_up(m * (2 * sg - 1) / (2 * nseg - 2), kTile);
    if (annot) {
        HIPC(ws->annot.ensure((size_t)m * K * 8));
        const int dev = panel->device;
        const uint64_t r1 = seg_r[1];
        upload = std::thread([ws, annot, m, K, dev, nseg, r1, &upload_err] {
            hipError_t e = hipSetDevice(dev);
            if (e == hipSuccess)
                e = nseg > 1 ? upload_annot_rows(ws, annot, m, K, 0, r1, ws->up_stream)
                             : ring_upload(ws, ws->annot.p, annot, (size_t)m * K * 8, ws->up_stream);
            if (e == hipSuccess) e = hipEventRecord(ws->ev[8], ws->up_stream);
            upload_err = e;
        });
    }
    // the MAF goes back to the host on a second stream while the band runs
    const bool has_pq = !std::isnan(cfg->pq_exp);
    const bool maf_aware = d && cfg->sketch_maf_aware;
    const bool early_stats = has_pq || maf_aware;
    const bool early_maf = dev_maf && maf_out && !early_stats;
    c.early_maf = early_maf;
    c.early_stats = early_stats;
    if (early_maf) {
        HIPC(ws->pinned_maf.ensure((size_t)m * 8));
        HIPC(hipEventRecord(ws->ev[6], st));
        HIPC(hipStreamWaitEvent(ws->copy_stream, ws->ev[6], 0));
        HIPC(hipMemcpyAsync(ws->pinned_maf.p, ws->mafd.p, m * 8, hipMemcpyDeviceToHost,
                            ws->copy_stream));
        HIPC(hipEventRecord(ws->ev[7], ws->copy_stream));
    }
    HIPC(hipEventRecord(ws->ev[1], st));
    const uint32_t nkb = (uint32_t)(round_up(n, kKBlock) / kKBlock);
    if (fused) {
        HIPC(launch_sketch_fused_scale(sp, d_musc, m_pad, (double)n, ws->sk_norm.as<double>(), st));
    } else if (d) {
        HIPC(ws->sk_norm.ensure(sketch_norm_bytes(sp, m_pad)));
        HIPC(launch_sketch(sp, sk8 ? kSketchI8 : kSketchF32, ws->tiled.as<uint32_t>(), nkb,
                           d_musc, (uint32_t)m, m_pad, (double)n, ws->sk_norm.as<double>(),
                           ws->sk_cols.p, nullptr, nullptr, st));
    }
    if (d) {
        if (sk8) {  // the range flag comes back for l2_band's choice
            HIPC(ws->pinned_flag.ensure(16));
            HIPC(hipMemcpyAsync(ws->pinned_flag.p,
                                sketch_i8_flag(sp, ws->sk_norm.as<double>(), m_pad), 4,
                                hipMemcpyDeviceToHost, st));
            HIPC(hipEventRecord(ws->ev[9], st));
        }
    }
    HIPC(hipEventRecord(ws->ev[2], st));

    // 2. host: pair bounds and the band's tile list, staged in a pinned
    //    buffer of their own while the device tiles
    std::vector<uint64_t>& lo = ws->host.lo;
    bool any_full = false;
    pair_bounds(snps, m, cfg, lo, &any_full, &ws->host.coord);
    c.any_full = any_full;
    if (!cfg->yes_really && any_full)
        printf("WARNING: LD window spans the entire chromosome. "
               "Use --yes-really to silence this warning.\n");
    BandPlan& plan = ws->host.plan;
    plan_band(lo, m, plan);
    const size_t off_lo = 0, off_tiles = round_up((size_t)m_pad * 4, 16);
    const size_t off_j0 = off_tiles + round_up((size_t)plan.ntiles * 8, 16);
    const size_t off_ts = off_j0 + round_up((size_t)plan.nb * 4, 16);
    const size_t off_st = off_ts + round_up((size_t)(plan.nb + 1) * 4, 16);
    const size_t off_maf = off_st + (size_t)m_pad * 16;
    const size_t off_l2 = off_maf + (size_t)m_pad * 8;
    // results come back through pinned memory: with the call's metadata
    // when small, else through the pinned ring
    const bool stage_l2 = (size_t)m * K * 8 <= (8u << 20);
    c.stage_l2 = stage_l2;
    c.off_st = off_st;
    c.off_l2 = off_l2;
    const size_t meta_bytes = off_l2 + (stage_l2 ? (size_t)m * K * 8 : 0);
    // nothing in flight reads pinned_meta (the previous call ended with a
    // synchronize), so it may grow here
    HIPC(ws->pinned_meta.ensure(meta_bytes));
    char* hp = ws->pinned_meta.as<char>();
    uint32_t* h_lo = reinterpret_cast<uint32_t*>(hp + off_lo);
    {
        const int T = task_count(m_pad);
        std::vector<uint64_t> part(T, 0);
        parallel_tasks(T, [&](int t) {
            const uint64_t a = (uint64_t)m_pad * t / T, b = (uint64_t)m_pad * (t + 1) / T;
            uint64_t pairs = 0;
            for (uint64_t j = a; j < b; ++j) {
                const uint64_t l = j < m ? lo[j] : j;
                h_lo[j] = (uint32_t)l;
                pairs += j - l;
            }
            part[t] = pairs;
        });
        plan.pairs_included = 0;
        for (uint64_t v : part) plan.pairs_included += v;
    }
    fill_band_tiles(plan, reinterpret_cast<int2*>(hp + off_tiles));
    std::memcpy(hp + off_j0, plan.j0.data(), (size_t)plan.nb * 4);
    std::memcpy(hp + off_ts, plan.tstart.data(), (size_t)(plan.nb + 1) * 4);
    const double t_prep = now_s();
    c.t_prep = t_prep;

    HIPC(ws->lo.ensure((size_t)m_pad * 4));
    const uint32_t Kp = aeff_stride((uint32_t)K);
    HIPC(ws->aeff.ensure((size_t)m_pad * Kp * 8));
    HIPC(ws->tiles.ensure((size_t)plan.ntiles * 8));
    HIPC(ws->j0.ensure((size_t)plan.nb * 4));
    HIPC(ws->tstart.ensure((size_t)(plan.nb + 1) * 4));
    HIPC(ws->rowp.ensure((size_t)plan.ntiles * K * kTile * 8));
    HIPC(ws->colp.ensure((size_t)plan.ntiles * K * kTile * 8));
    HIPC(ws->l2.ensure((size_t)m * K * 8));
    // the plan goes up on a stream of its own, beside the tiling still
    // running on st; st waits for it before anything that reads it
    hipStream_t pst = ws->plan_stream;
    HIPC(hipMemcpyAsync(ws->lo.p, h_lo, (size_t)m_pad * 4, hipMemcpyHostToDevice, pst));
    HIPC(hipMemcpyAsync(ws->tiles.p, hp + off_tiles, (size_t)plan.ntiles * 8,
                        hipMemcpyHostToDevice, pst));
    HIPC(hipMemcpyAsync(ws->j0.p, hp + off_j0, (size_t)plan.nb * 4, hipMemcpyHostToDevice, pst));
    HIPC(hipMemcpyAsync(ws->tstart.p, hp + off_ts, (size_t)(plan.nb + 1) * 4,
                        hipMemcpyHostToDevice, pst));
    HIPC(hipEventRecord(ws->plan_ev, pst));
    HIPC(hipStreamWaitEvent(st, ws->plan_ev, 0));

    // 3. MAF (and pq-scaled annotations, kappa) on the host from the exact
    //    stats; fetched now and waited for only when the device needs
    //    something derived from them.  Else, for the sketch form's host MAF,
    //    they come back on the copy stream while the band runs and l2_end
    //    forms the MAF before it waits for the result.
    uint32_t* h_stats = reinterpret_cast<uint32_t*>(hp + off_st);
    if (maf_out && !early_maf && !early_stats) {
        HIPC(hipStreamWaitEvent(ws->copy_stream, ws->ev[1], 0));
        HIPC(hipMemcpyAsync(h_stats, ws->stats.p, (size_t)m_pad * 16, hipMemcpyDeviceToHost,
                            ws->copy_stream));
        HIPC(hipEventRecord(ws->ev[7], ws->copy_stream));
        c.stats_fetched = true;
    }
    double* h_maf = reinterpret_cast<double*>(hp + off_maf);
    std::vector<double>& maf = ws->host.maf;
    if (early_stats) {
        if (dev_maf) {
            HIPC(hipMemcpyAsync(h_maf, ws->mafd.p, m * 8, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            maf.assign(h_maf, h_maf + m);
        } else {
            HIPC(hipMemcpyAsync(h_stats, ws->stats.p, (size_t)m_pad * 16, hipMemcpyDeviceToHost, st));
            HIPC(hipStreamSynchronize(st));
            compute_maf_host(c);
        }
    }
    if (maf_aware) {
        // per_snp_kappa (compute.rs:525-538) from the integer moments:
        // sum g^3 = 8 n2 + het, sum g^4 = 16 n2 + het, n2 = (sum - het) / 2
        std::vector<double>& kap = ws->host.aux;
        kap.assign(m_pad, 0.0);
        for (uint64_t j = 0; j < m; ++j) {
            const uint32_t S = h_stats[4 * j], C = h_stats[4 * j + 1], SS = h_stats[4 * j + 2],
                           H = h_stats[4 * j + 3];
            const uint32_t n2 = (S - H) / 2;
            const double cube = 8.0 * n2 + H, fourth = 16.0 * n2 + H;
            if (C == 0) continue;
            const double cf = (double)C, mean = (double)S / cf;
            // mean.powi(4) of the reference: (x*x)*(x*x)
            const double mean2 = mean * mean;
            const double m4 = fourth - 4.0 * mean * cube + 6.0 * mean * mean * (double)SS -
                              3.0 * (mean2 * mean2) * cf;
            const double css = (double)SS - cf * mean * mean;
            kap[j] = css <= 0.0 ? 0.0 : m4 / (css * css);
        }
        HIPC(hipMemcpyAsync(ws->maf.p, kap.data(), kap.size() * 8, hipMemcpyHostToDevice, st));
        HIPC(hipStreamSynchronize(st));
    }
    if (has_pq) {
        // (maf (1 - maf))^S on the host with libm pow, like the reference
        std::vector<double>& pq = ws->host.aux;
        pq.resize(m);
        for (uint64_t j = 0; j < m; ++j) pq[j] = std::pow(maf[j] * (1.0 - maf[j]), cfg->pq_exp);
        HIPC(ws->pq.ensure(m * 8));
        HIPC(hipMemcpyAsync(ws->pq.p, pq.data(), m * 8, hipMemcpyHostToDevice, st));
        HIPC(hipStreamSynchronize(st));  // `pq` is pageable and local
    }
    if (annot) {
        upload.join();
        HIPC(upload_err);
        HIPC(hipStreamWaitEvent(st, ws->ev[8], 0));
    }
    // A_eff of every row, or of the first segment's rows
    hipLaunchKernelGGL(k_aeff, dim3(seg_r[1] / 64, Kp / 16 + (Kp == 1)), dim3(256), 0, st,
                       ws->aeff.as<double>(), annot ? ws->annot.as<double>() : nullptr,
                       (uint64_t)(nseg > 1 ? seg_r[1] : m), 0u,
                       has_pq ? ws->pq.as<double>() : nullptr, (uint32_t)m, (uint32_t)K, Kp);
    HIPC(hipGetLastError());
    c.annot = annot;
    c.nseg = nseg;
    c.Kp = Kp;
    c.has_pq = has_pq;
    for (uint32_t sg = 0; sg <= nseg; ++sg) c.seg_I[sg] = seg_r[sg] / kTile;

    // 3. band contraction
    BandArgs ba{};
    ba.tiled = ws->tiled.as<u32x4>();
    ba.tiles = ws->tiles.as<int2>();
    ba.lo = ws->lo.as<uint32_t>();
    ba.cs = d_cs;
    ba.rs = d_rs;
    ba.lut = d_lut;
    ba.aeff = ws->aeff.as<double>();
    ba.rowp = ws->rowp.as<double>();
    ba.colp = ws->colp.as<double>();
    ba.sk = d ? ws->sk_cols.as<float>() : nullptr;
    ba.sk_scale = d && !sk8 ? sketch_scales(sp, ws->sk_norm.as<double>(), m_pad) : nullptr;
    if (sk8) {
        ba.skw = ws->sk_cols.as<u32x4>();
        ba.ske = ba.skw + (size_t)m_pad * sp->d_pad / 16;
        ba.sk8 = sketch_i8_consts(sp, ws->sk_norm.as<double>(), m_pad);
        ba.sk_tt = sp->tt;
    }
    ba.kappa = maf_aware ? ws->maf.as<double>() : nullptr;
    ba.nkb = d ? (uint32_t)sp->nblocks : nkb;
    ba.ntiles = plan.ntiles;
    ba.per_xcd = (plan.ntiles + 7) / 8;
    ba.m = (uint32_t)m;
    ba.m_pad = m_pad;
    ba.K = (uint32_t)K;
    const double nd = (double)n, n_inv = 1.0 / nd;
    const double denom = n > 2 ? nd - 2.0 : nd;
    ba.r2u_a = 1.0 + 1.0 / denom;
    ba.r2u_b = -1.0 / denom;
    ba.n_inv_sq_r2u_a = n_inv * n_inv * ba.r2u_a;
    ba.n_inv_sq = n_inv * n_inv;
    ba.d_f = (double)d;
    ba.n_pad = (double)round_up(n, kKBlock);

    c.ba = ba;
    c.sketch_plan = sp;
    // everything the band reads is queued on the stream before this event
    HIPC(hipEventRecord(ws->ev[3], st));
    c.done = false;
    return LDSC_OK;
}

static hipError_t launch_band_form(int form, const BandArgs& ba, hipStream_t st) {
    if (form == kBandSketch) return launch_band_sk(ba, st);
    if (form == kBandSketch8) return launch_band_sk8(ba, st);
    if (form == kBandF32)
        hipLaunchKernelGGL(k_band_f32_2w, dim3(8 * ba.per_xcd), dim3(128), 0, st, ba);
    else if (form == kBandFp4)
        hipLaunchKernelGGL(k_band_fp4, dim3(8 * ba.per_xcd), dim3(64), 0, st, ba);
    else
        hipLaunchKernelGGL(k_band_i8, dim3(8 * ba.per_xcd), dim3(128), 0, st, ba);
    return hipGetLastError();
}

int l2_band(L2Call* call, char* err, uint64_t err_len) {
    L2Call& c = *call;
    Workspace* ws = c.ws;
    hipStream_t st = ws->stream;
    c.nprog = 0;
    if (c.form == kBandSketch8) {
        // a W or E value outside [-127, 127]: the bucket sums again as f32
        // values (from the same tiles) and the f32 contraction
        HIPC(hipEventSynchronize(ws->ev[9]));
        if (*ws->pinned_flag.as<int32_t>()) {
            const SketchPlan* sp = c.sketch_plan;
            const uint32_t nkb = (uint32_t)(round_up(c.panel->n, kKBlock) / kKBlock);
            if (c.sk_fused) {  // the one-pass sums left no tiled rows
                const bool contiguous = c.bed_contiguous;
                uint32_t parts = 0;
                const int rc = gather_tile_chunks(
                    c.panel, contiguous ? nullptr : ws->bed_idx.as<uint64_t>(),
                    contiguous ? c.snps[0].bed_idx : 0, c.m, sp->d_local, (uint32_t)sp->chunk,
                    c.m_pad, ws, &parts, err, err_len);
                if (rc) return rc;
            }
            const float2* d_musc =
                reinterpret_cast<const float2*>(ws->info.as<char>() + (size_t)c.m_pad * 32);
            HIPC(launch_sketch(sp, kSketchF32, ws->tiled.as<uint32_t>(), nkb, d_musc,
                               (uint32_t)c.m, c.m_pad, (double)c.panel->n,
                               ws->sk_norm.as<double>(), ws->sk_cols.p, nullptr, nullptr, st));
            c.ba.sk_scale = sketch_scales(sp, ws->sk_norm.as<double>(), c.m_pad);
            c.form = kBandSketch;
        }
    }
    const std::vector<int>& ts = ws->host.plan.tstart;
    const std::vector<int>& j0v = ws->host.plan.j0;
    const uint32_t nb = ws->host.plan.nb;
    if (c.nseg <= 1 && c.cb && nb > 1) {
        // With a progress callback the band runs as consecutive pieces of
        // block rows, an event after each.  Once the pieces up to block row
        // I are done, the SNPs below 64 j0[I] are complete: their row tiles
        // ran, and no later block row reaches their column block.  The tiles
        // and the per-tile arithmetic are those of one launch.
        const uint32_t P = std::min<uint32_t>(Workspace::kProgPieces, nb);
        for (uint32_t p = 0; p < P; ++p) {
            const uint32_t I0 = (uint32_t)((uint64_t)nb * p / P);
            const uint32_t I1 = (uint32_t)((uint64_t)nb * (p + 1) / P);
            BandArgs b = c.ba;
            b.tile0 = (uint32_t)ts[I0];
            b.ntiles = (uint32_t)ts[I1] - b.tile0;
            b.per_xcd = (b.ntiles + 7) / 8;
            if (b.ntiles) HIPC(launch_band_form(c.form, b, st));
            HIPC(hipEventRecord(ws->prog_ev[p], st));
            c.prog_ev[p] = ws->prog_ev[p];
            c.prog_snps[p] = I1 < nb ? std::min<uint32_t>((uint32_t)c.m, (uint32_t)j0v[I1] * kTile)
                                     : (uint32_t)c.m;
        }
        c.nprog = P;
        HIPC(hipEventRecord(ws->ev[4], st));
        return LDSC_OK;
    }
    if (c.nseg <= 1) {
        HIPC(launch_band_form(c.form, c.ba, st));
    } else {
        // segment sg: its annotation rows go up (the first already did) while
        // the band of the segments before it runs, then its A_eff rows, then
        // the band of its row blocks, whose columns lie in this segment or
        // the ones before
        uint32_t red_to = 0;            // rows reduced so far
        uint32_t dl_a = 0, dl_b = 0;    // rows reduced after the last band, not yet back
        for (uint32_t sg = 0; sg < c.nseg; ++sg) {
            const uint32_t r0 = c.seg_I[sg] * kTile, r1 = c.seg_I[sg + 1] * kTile;
            if (sg > 0) {
                const uint64_t u1 = std::min<uint64_t>(r1, c.m);
                HIPC(upload_annot_rows(ws, c.annot, c.m, c.K, r0, u1, ws->up_stream));
                HIPC(hipEventRecord(ws->ev[8], ws->up_stream));
                HIPC(hipStreamWaitEvent(st, ws->ev[8], 0));
                hipLaunchKernelGGL(k_aeff, dim3((r1 - r0) / 64, c.Kp / 16 + (c.Kp == 1)), dim3(256),
                                   0, st, ws->aeff.as<double>(),
                                   ws->annot.as<double>() + (size_t)r0 * c.K, (uint64_t)(u1 - r0), r0,
                                   c.has_pq ? ws->pq.as<double>() : nullptr, (uint32_t)c.m,
                                   (uint32_t)c.K, c.Kp);
                HIPC(hipGetLastError());
            }
            BandArgs b = c.ba;
            b.tile0 = (uint32_t)ts[c.seg_I[sg]];
            b.ntiles = (uint32_t)ts[c.seg_I[sg + 1]] - b.tile0;
            b.per_xcd = (b.ntiles + 7) / 8;
            if (b.ntiles) HIPC(launch_band_form(c.form, b, st));
            // rows this band completes (their row tiles are in this or earlier
            // segments, and no later row block reaches their column block) are
            // reduced right after it, on the band's stream
            const uint32_t done_to =
                sg + 1 < c.nseg
                    ? std::min<uint32_t>((uint32_t)c.m, (uint32_t)j0v[c.seg_I[sg + 1]] * kTile)
                    : (uint32_t)c.m;
            const uint32_t prev_a = dl_a, prev_b = dl_b;
            dl_a = dl_b = red_to;
            if (sg + 1 == c.nseg) HIPC(hipEventRecord(ws->ev[4], st));
            if (done_to > red_to) {
                const uint64_t nred = (uint64_t)(done_to - red_to) * c.K;
                hipLaunchKernelGGL(k_reduce, dim3((unsigned)((nred + 255) / 256)), dim3(256), 0, st,
                                   ws->rowp.as<double>(), ws->colp.as<double>(),
                                   ws->aeff.as<double>(), ws->j0.as<int>(), ws->tstart.as<int>(),
                                   ws->host.plan.nb, red_to, done_to - red_to, (uint32_t)c.K, 0u,
                                   (uint32_t)c.K, ws->l2.as<double>());
                HIPC(hipGetLastError());
                dl_b = done_to;
                red_to = done_to;
            }
            HIPC(hipEventRecord(ws->red_ev[sg], st));
            c.prog_ev[sg] = ws->red_ev[sg];
            c.prog_snps[sg] = done_to;
            c.nprog = sg + 1;
            // the rows reduced after the previous segment's band go back while
            // this band runs
            if (prev_b > prev_a) {
                HIPC(hipStreamWaitEvent(ws->copy_stream, ws->red_ev[sg - 1], 0));
                HIPC(download_l2_rows(ws, c.l2_out, c.m, c.K, prev_a, prev_b, ws->copy_stream));
            }
        }
        // the last reduced rows go back in l2_end, after the last reduce
        c.seg_dl_a = dl_a;
        c.seg_dl_b = dl_b;
        HIPC(hipEventRecord(ws->ev[5], st));
        HIPC(hipStreamWaitEvent(ws->copy_stream, ws->ev[5], 0));
        return LDSC_OK;
    }
    HIPC(hipEventRecord(ws->ev[4], st));
    return LDSC_OK;
}

int band_batch_launch(L2Call* const* calls, uint32_t ncalls, BandBatchSlot* slot, hipStream_t bst,
                      char* err, uint64_t err_len) {
    if (ncalls == 0) return LDSC_OK;
    if (ncalls > BandBatchSlot::kMaxCalls)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG, "band batch of %u calls is too large",
                    ncalls);
    const int form = calls[0]->form;
    // the slot's previous launch has read its table
    if (slot->used) HIPC(hipEventSynchronize(slot->done));
    if (!slot->d_args) HIPC(hipMalloc(&slot->d_args, BandBatchSlot::kBytes));
    if (!slot->h_args) HIPC(hipHostMalloc(&slot->h_args, BandBatchSlot::kBytes, hipHostMallocDefault));
    if (!slot->done) HIPC(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming));
    BandArgs* h_args = reinterpret_cast<BandArgs*>(slot->h_args);
    uint32_t* h_starts = reinterpret_cast<uint32_t*>(h_args + BandBatchSlot::kMaxCalls);
    uint32_t total = 0;
    for (uint32_t i = 0; i < ncalls; ++i) {
        if (calls[i]->form != form || calls[i]->panel->device != calls[0]->panel->device)
            return fail(err, err_len, LDSC_ERR_INVALID_ARG, "mixed band batch");
        h_args[i] = calls[i]->ba;
        h_starts[i] = total;
        total += calls[i]->ba.ntiles;
        // the band reads what the call queued before its ev[3]
        HIPC(hipStreamWaitEvent(bst, calls[i]->ws->ev[3], 0));
    }
    h_starts[ncalls] = total;
    BandBatch b{};
    b.args = reinterpret_cast<const BandArgs*>(slot->d_args);
    b.starts = reinterpret_cast<const uint32_t*>(reinterpret_cast<const BandArgs*>(slot->d_args) +
                                                 BandBatchSlot::kMaxCalls);
    b.nreq = ncalls;
    b.ntiles = total;
    b.per_xcd = (total + 7) / 8;
    HIPC(hipMemcpyAsync(slot->d_args, slot->h_args, BandBatchSlot::kBytes, hipMemcpyHostToDevice,
                        bst));
    if (total > 0) {
        if (form == kBandF32)
            hipLaunchKernelGGL(k_band_f32_2w_batch, dim3(8 * b.per_xcd), dim3(128), 0, bst, b);
        else if (form == kBandFp4)
            hipLaunchKernelGGL(k_band_fp4_batch, dim3(8 * b.per_xcd), dim3(64), 0, bst, b);
        else if (form == kBandI8)
            hipLaunchKernelGGL(k_band_i8_batch, dim3(8 * b.per_xcd), dim3(128), 0, bst, b);
        else
            return fail(err, err_len, LDSC_ERR_INVALID_ARG, "this band form is not batched");
        HIPC(hipGetLastError());
    }
    HIPC(hipEventRecord(slot->done, bst));
    slot->used = true;
    for (uint32_t i = 0; i < ncalls; ++i) HIPC(hipEventRecord(calls[i]->ws->ev[4], bst));
    return LDSC_OK;
}

// Ends of the reference's chunks over the call's SNPs: one grid over all of
// them (compute.rs:1281-1283), or one per chromosome run when the call
// stands for run()'s per-chromosome calls (mod.rs:478-553).  Their number is
// L2Progress::chunks_total.
static void chunk_ends(const L2Call& c, std::vector<uint64_t>& ends) {
    ends.clear();
    const uint64_t m = c.m, cs = c.cfg->chunk_size;
    for (uint64_t r0 = 0; r0 < m;) {
        uint64_t r1 = m;
        if (c.cfg->per_chromosome) {
            r1 = r0 + 1;
            while (r1 < m && c.snps[r1].chr == c.snps[r0].chr) ++r1;
        }
        for (uint64_t s = r0; s < r1; s += cs) ends.push_back(std::min(s + cs, r1));
        r0 = r1;
    }
}

int l2_end(L2Call* call, char* err, uint64_t err_len) {
    L2Call& c = *call;
    Workspace* ws = c.ws;
    hipStream_t st = ws->stream;
    const uint64_t m = c.m, K = c.K;
    const bool stage_l2 = c.stage_l2, early_maf = c.early_maf;
    double* l2_out = c.l2_out;
    double* maf_out = c.maf_out;
    ldsc_l2_perf* perf = c.perf;
    BandPlan& plan = ws->host.plan;
    char* hp = ws->pinned_meta.as<char>();
    std::vector<double>& maf = ws->host.maf;
    // the band ran on this stream or on a batch stream that recorded ev[4]
    HIPC(hipStreamWaitEvent(st, ws->ev[4], 0));

    // 4. deterministic reduction, in groups of annotations when the result
    //    goes back through the ring, so that its copy starts with the first
    //    (a segmented band reduced its rows segment by segment already)
    const bool seg = c.nseg > 1;
    const uint32_t ngroups =
        seg ? 0u : stage_l2 ? 1u : (uint32_t)std::min<uint64_t>(K, Workspace::kRedGroups);
    for (uint32_t g = 0; g < ngroups; ++g) {
        const uint32_t k0 = (uint32_t)(K * g / ngroups), k1 = (uint32_t)(K * (g + 1) / ngroups);
        const uint64_t nred = m * (k1 - k0);
        hipLaunchKernelGGL(k_reduce, dim3((unsigned)((nred + 255) / 256)), dim3(256), 0, st,
                           ws->rowp.as<double>(), ws->colp.as<double>(), ws->aeff.as<double>(),
                           ws->j0.as<int>(), ws->tstart.as<int>(), plan.nb, 0u, (uint32_t)m,
                           (uint32_t)K, k0, k1 - k0, ws->l2.as<double>());
        HIPC(hipGetLastError());
        if (!stage_l2) HIPC(hipEventRecord(ws->red_ev[g], st));
    }
    if (!seg) HIPC(hipEventRecord(ws->ev[5], st));
    // progress while the band runs: one call per progress point that
    // completes a chunk not reported yet (the last call, with every chunk
    // done, comes once the results are back)
    std::vector<uint64_t>& ends = ws->host.chunk_end;
    double t_cb = c.t_start;
    if (c.cb) {
        chunk_ends(c, ends);
        uint64_t reported = 0;
        for (uint32_t p = 0; p < c.nprog; ++p) {
            HIPC(hipEventSynchronize(c.prog_ev[p]));
            const uint64_t done =
                (uint64_t)(std::upper_bound(ends.begin(), ends.end(), (uint64_t)c.prog_snps[p]) -
                           ends.begin());
            if (done <= reported || done >= ends.size()) continue;
            const double t = now_s();
            ldsc_l2_progress pr{done, ends.size(), ends[done - 1], m, (t - t_cb) * 1e3};
            c.cb(&pr, c.cb_ctx);
            reported = done;
            t_cb = t;
        }
    }
    if (early_maf) {  // copy it out while the band kernel runs
        HIPC(hipEventSynchronize(ws->ev[7]));
        std::memcpy(maf_out, ws->pinned_maf.p, m * 8);
    } else if (c.stats_fetched) {  // the sketch form's MAF, while the band runs
        HIPC(hipEventSynchronize(ws->ev[7]));
        compute_maf_host(c);
    }
    double* h_l2 = reinterpret_cast<double*>(hp + c.off_l2);
    if (stage_l2) {
        HIPC(hipMemcpyAsync(h_l2, ws->l2.p, (size_t)m * K * 8, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
        std::memcpy(l2_out, h_l2, (size_t)m * K * 8);
    } else if (seg) {
        // the rows of the last segment's reduce (queued after it on the copy stream)
        if (c.seg_dl_b > c.seg_dl_a)
            HIPC(download_l2_rows(ws, l2_out, m, K, c.seg_dl_a, c.seg_dl_b, ws->copy_stream));
        HIPC(hipStreamSynchronize(st));
    } else {
        // each piece of the result is copied once the group of annotations
        // holding its last byte is reduced
        uint32_t waited = 0;
        auto ready = [&](size_t end) -> hipError_t {
            const uint64_t klast = (end - 1) / ((uint64_t)m * 8);
            uint32_t g = 0;
            while ((uint64_t)K * (g + 1) / ngroups <= klast) ++g;
            for (; waited <= g; ++waited) {
                hipError_t e = hipStreamWaitEvent(ws->copy_stream, ws->red_ev[waited], 0);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        };
        HIPC(ring_download(ws, l2_out, ws->l2.p, (size_t)m * K * 8, ws->copy_stream, ready));
        HIPC(hipStreamSynchronize(st));
    }
    if (maf_out && !early_maf) std::memcpy(maf_out, maf.data(), m * 8);

    if (perf) {
        float t01 = 0, t12 = 0, t34 = 0, t45 = 0;
        (void)hipEventElapsedTime(&t01, ws->ev[0], ws->ev[1]);
        (void)hipEventElapsedTime(&t12, ws->ev[1], ws->ev[2]);
        (void)hipEventElapsedTime(&t34, ws->ev[3], ws->ev[4]);
        (void)hipEventElapsedTime(&t45, ws->ev[4], ws->ev[5]);
        std::memset(perf, 0, sizeof(*perf));
        perf->m = m;
        perf->n_indiv = c.panel->n;
        perf->bed_read_secs = c.t_prep - c.t_start;
        perf->norm_secs = t01 * 1e-3;
        perf->sketch_secs = t12 * 1e-3;
        perf->bb_dot_secs = t34 * 1e-3;
        perf->ring_store_secs = t45 * 1e-3;
        perf->total_secs = now_s() - c.t_start;
        perf->pairs_included = plan.pairs_included;
        perf->pairs_evaluated = (uint64_t)plan.ntiles * kTile * kTile;
        perf->tiles = plan.ntiles;
        perf->any_full_chr_window = c.any_full ? 1 : 0;
        perf->sketch_int8 = c.form == kBandSketch8 ? 1 : 0;
    }
    if (c.cb) {
        ldsc_l2_progress pr{ends.size(), ends.size(), m, m, (now_s() - t_cb) * 1e3};
        c.cb(&pr, c.cb_ctx);
    }
    c.done = true;
    return LDSC_OK;
}

int run_l2(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
           const double* annot, uint64_t K, double* l2_out, double* maf_out, ldsc_l2_perf* perf,
           ldsc_progress_cb cb, void* cb_ctx, Workspace* ws, char* err, uint64_t err_len) {
    L2Call c;
    int rc = l2_begin(panel, snps, m, cfg, annot, K, l2_out, maf_out, perf, cb, cb_ctx, ws, &c,
                      err, err_len, /*segment_annot=*/true);
    if (rc || c.done) return rc;
    if ((rc = l2_band(&c, err, err_len))) return rc;
    return l2_end(&c, err, err_len);
}

}  // namespace ldsc

using namespace ldsc;

// ------------------------------------------------------------------------
// C ABI: panels
// ------------------------------------------------------------------------
// Moves `rows` BED rows of `bps` bytes, packed back to back in `src`, to
// rows `stride` bytes apart in `dst` (stride = bps rounded up to 16), with
// the padding bytes zero.  One thread writes one 4-byte word of the output.
__global__ void k_restride_rows(const uint8_t* __restrict__ src, uint64_t bps, uint64_t rows,
                                uint8_t* __restrict__ dst, uint64_t stride) {
    const uint64_t wpr = stride / 4;
    const uint64_t total = rows * wpr;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = t / wpr, w = t - r * wpr;
        const uint8_t* row = src + r * bps;
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint64_t off = 4 * w + b;
            if (off < bps) v |= (uint32_t)row[off] << (8 * b);
        }
        reinterpret_cast<uint32_t*>(dst + r * stride)[w] = v;
    }
}

// Uploads the BED body (m rows of bps bytes after the 3-byte header) into
// the panel's strided rows.  A pageable copy reaches PCIe speed only when it
// is large and contiguous, so the body goes up in contiguous pieces of at
// most kUploadPiece bytes into a device staging buffer, and a kernel moves
// each piece's rows into place.  When bps is already a multiple of 16 the
// pieces land in place directly.
static hipError_t upload_rows(ldsc_panel* p, const uint8_t* body) {
    constexpr uint64_t kUploadPiece = 256ull << 20;
    const uint64_t bps = p->bps, stride = p->stride, m = p->m_bim;
    const uint64_t piece_rows = std::max<uint64_t>(1, kUploadPiece / bps);
    hipStream_t st = nullptr;
    uint8_t* stage = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess && bps != stride)
        e = hipMalloc(&stage, std::min(piece_rows, m) * bps);
    for (uint64_t r0 = 0; e == hipSuccess && r0 < m; r0 += piece_rows) {
        const uint64_t rows = std::min(piece_rows, m - r0);
        if (bps == stride) {
            e = hipMemcpyAsync(p->d_slabs + r0 * stride, body + r0 * bps, rows * bps,
                               hipMemcpyHostToDevice, st);
            continue;
        }
        e = hipMemcpyAsync(stage, body + r0 * bps, rows * bps, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) break;
        const uint64_t words = rows * (stride / 4);
        const uint32_t grid = (uint32_t)std::min<uint64_t>((words + 255) / 256, 8192);
        hipLaunchKernelGGL(k_restride_rows, dim3(grid), dim3(256), 0, st, stage, bps, rows,
                           p->d_slabs + r0 * stride, stride);
        e = hipGetLastError();
    }
    if (st) {
        const hipError_t es = hipStreamSynchronize(st);
        if (e == hipSuccess) e = es;
        (void)hipStreamDestroy(st);
    }
    if (stage) (void)hipFree(stage);
    return e;
}

extern "C" int ldsc_l2_open_panel(const uint8_t* bed, uint64_t bed_len, uint64_t n_fam,
                                  uint64_t m_bim, const int64_t* keep_idx, uint64_t n_keep,
                                  ldsc_panel** out, char* err, uint64_t err_len) {
    if (!out) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!bed && bed_len) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "bed bytes are NULL");
    // Bed::from_source validation (src/bed.rs:76-109)
    if (bed_len < 3) return fail(err, err_len, LDSC_ERR_BAD_BED, "BED too short (%lluB)", (unsigned long long)bed_len);
    if (bed[0] != 0x6c || bed[1] != 0x1b) return fail(err, err_len, LDSC_ERR_BAD_BED, "BED has invalid magic bytes");
    if (bed[2] != 0x01) return fail(err, err_len, LDSC_ERR_BAD_BED, "BED is not SNP-major (mode=1 required)");
    const uint64_t bps = (n_fam + 3) / 4;
    const uint64_t expect = bps * m_bim + 3;
    if (bed_len != expect)
        return fail(err, err_len, LDSC_ERR_BAD_BED, "BED has invalid length (expected %llu, got %llu)",
                    (unsigned long long)expect, (unsigned long long)bed_len);
    // compute_l2_from_bytes (src/l2/mod.rs:1062-1069)
    if (m_bim == 0) return fail(err, err_len, LDSC_ERR_EMPTY, "BIM has zero SNPs");
    if (n_fam == 0) return fail(err, err_len, LDSC_ERR_EMPTY, "FAM has zero individuals");
    if (keep_idx && n_keep == 0) return fail(err, err_len, LDSC_ERR_EMPTY, "--keep selects zero individuals");
    const uint64_t n = keep_idx ? n_keep : n_fam;
    if (n >= (1ull << 31)) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "too many individuals");
    if (keep_idx)
        for (uint64_t i = 0; i < n_keep; ++i)
            if (keep_idx[i] < 0 || (uint64_t)keep_idx[i] >= n_fam)
                return fail(err, err_len, LDSC_ERR_INVALID_ARG, "--keep index %lld out of range",
                            (long long)keep_idx[i]);
    auto* p = new ldsc_panel();
    p->n_fam = n_fam;
    p->m_bim = m_bim;
    p->bps = bps;
    p->stride = round_up(bps, 16);
    p->n = n;
    // rows 16-byte aligned and padded to whole dwords, so that device reads
    // of a row are aligned vector loads that never leave the allocation
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc(&p->d_slabs, p->stride * m_bim);
    if (e == hipSuccess) e = upload_rows(p, bed + 3);
    if (e == hipSuccess && keep_idx) {
        p->keep.assign(keep_idx, keep_idx + n_keep);
        e = hipMalloc(&p->d_keep, n_keep * 8);
        if (e == hipSuccess) e = hipMemcpy(p->d_keep, keep_idx, n_keep * 8, hipMemcpyHostToDevice);
        const std::vector<int32_t> k32 =
            lane_major_map(std::vector<int32_t>(keep_idx, keep_idx + n_keep));
        if (e == hipSuccess) e = hipMalloc(&p->d_keep32, k32.size() * 4);
        if (e == hipSuccess) e = hipMemcpy(p->d_keep32, k32.data(), k32.size() * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        ldsc_l2_close_panel(p);
        return fail(err, err_len, e == hipErrorOutOfMemory ? LDSC_ERR_OOM : LDSC_ERR_DEVICE,
                    "uploading the panel failed: %s", hipGetErrorString(e));
    }
    *out = p;
    return LDSC_OK;
}

extern "C" int ldsc_l2_close_panel(ldsc_panel* p) {
    if (!p) return LDSC_OK;
    if (p->d_slabs) (void)hipFree(p->d_slabs);
    if (p->d_keep) (void)hipFree(p->d_keep);
    if (p->d_keep32) (void)hipFree(p->d_keep32);
    for (SketchPlan* sp : p->sketch_plans) {
        sp->release();
        delete sp;
    }
    delete p;
    return LDSC_OK;
}

extern "C" uint64_t ldsc_l2_panel_n_indiv(const ldsc_panel* p) { return p ? p->n : 0; }

extern "C" int ldsc_l2_compute(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m,
                               const ldsc_l2_config* cfg, const double* annot, uint64_t K,
                               double* l2_out, double* maf_out, ldsc_l2_perf* perf,
                               ldsc_progress_cb cb, void* cb_ctx, char* err, uint64_t err_len) {
    if (!panel) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "panel is NULL");
    Workspace* ws = acquire_workspace(panel->device);
    if (!ws) return fail(err, err_len, LDSC_ERR_DEVICE, "could not create a device workspace");
    int rc = run_l2(panel, snps, m, cfg, annot, K, l2_out, maf_out, perf, cb, cb_ctx, ws, err, err_len);
    release_workspace(ws);
    return rc;
}

extern "C" int ldsc_l2_decode(const ldsc_panel* panel, const uint64_t* bed_idx, uint64_t m,
                              int8_t* dosage, uint32_t* miss, uint32_t* stats, float* std_f32,
                              char* err, uint64_t err_len) {
    if (!panel) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "panel is NULL");
    if (m == 0) return LDSC_OK;
    if (!bed_idx) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "bed_idx is NULL");
    for (uint64_t j = 0; j < m; ++j)
        if (bed_idx[j] >= panel->m_bim)
            return fail(err, err_len, LDSC_ERR_INVALID_ARG, "bed_idx %llu out of range",
                        (unsigned long long)bed_idx[j]);
    if (m > 65535) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "decode at most 65535 SNPs per call");
    Workspace* ws = acquire_workspace(panel->device);
    if (!ws) return fail(err, err_len, LDSC_ERR_DEVICE, "could not create a device workspace");
    const uint64_t n = panel->n;
    const uint32_t words = (uint32_t)((n + 31) / 32);
    int rc = LDSC_OK;
    do {
        hipStream_t st = ws->stream;
        hipError_t e;
#define DCHK(x) if ((e = (x)) != hipSuccess) { rc = fail(err, err_len, LDSC_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e)); break; }
        DCHK(hipSetDevice(panel->device));
        DCHK(ws->bed_idx.ensure(m * 8));
        DCHK(ws->stats.ensure(m * 16));
        DCHK(ws->info.ensure(m * 32));
        DCHK(ws->scratch.ensure((dosage ? n * m : 0) + (miss ? (size_t)words * 4 * m : 0) +
                                (std_f32 ? n * m * 4 : 0) + 256));
        DCHK(hipMemcpyAsync(ws->bed_idx.p, bed_idx, m * 8, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_stats_only, dim3((unsigned)m), dim3(256), 0, st, panel->d_slabs,
                           panel->stride, panel->d_keep, (uint32_t)n, ws->bed_idx.as<uint64_t>(),
                           ws->stats.as<uint32_t>());
        DCHK(hipGetLastError());
        int2* d_cs = ws->info.as<int2>();
        double* d_rs = reinterpret_cast<double*>(ws->info.as<char>() + m * 8);
        float4* d_lut = reinterpret_cast<float4*>(ws->info.as<char>() + m * 16);
        hipLaunchKernelGGL(k_snp_info, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                           ws->stats.as<uint32_t>(), 1u, ws->stats.as<uint32_t>(), (uint32_t)m,
                           (uint32_t)n, d_cs, d_rs, d_lut, nullptr, nullptr, 0);
        DCHK(hipGetLastError());
        char* sp = ws->scratch.as<char>();
        int8_t* d_dos = dosage ? reinterpret_cast<int8_t*>(sp) : nullptr;
        sp += dosage ? n * m : 0;
        sp = reinterpret_cast<char*>(round_up((uintptr_t)sp, 16));
        uint32_t* d_miss = miss ? reinterpret_cast<uint32_t*>(sp) : nullptr;
        sp += miss ? (size_t)words * 4 * m : 0;
        sp = reinterpret_cast<char*>(round_up((uintptr_t)sp, 16));
        float* d_std = std_f32 ? reinterpret_cast<float*>(sp) : nullptr;
        if (dosage || miss || std_f32) {
            const uint32_t gx = (uint32_t)((round_up(n, 64)) / 64);
            hipLaunchKernelGGL(k_decode, dim3(gx, (unsigned)m), dim3(64), 0, st, panel->d_slabs,
                               panel->stride, panel->d_keep, (uint32_t)n, ws->bed_idx.as<uint64_t>(),
                               (uint32_t)m, d_lut, d_dos, d_miss, d_std);
            DCHK(hipGetLastError());
        }
        if (stats) DCHK(hipMemcpyAsync(stats, ws->stats.p, m * 16, hipMemcpyDeviceToHost, st));
        if (dosage) DCHK(hipMemcpyAsync(dosage, d_dos, n * m, hipMemcpyDeviceToHost, st));
        if (miss) DCHK(hipMemcpyAsync(miss, d_miss, (size_t)words * 4 * m, hipMemcpyDeviceToHost, st));
        if (std_f32) DCHK(hipMemcpyAsync(std_f32, d_std, n * m * 4, hipMemcpyDeviceToHost, st));
        DCHK(hipStreamSynchronize(st));
#undef DCHK
    } while (0);
    release_workspace(ws);
    return rc;
}
