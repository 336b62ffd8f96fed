// sketch.hip -- the CountSketch path (--sketch d).
//
// Reference: CountSketchProj (src/l2/compute.rs:171-302), the fused
// decode -> normalize -> signed scatter-add (countsketch_fused_project_f32,
// compute.rs:601-770; scatter_one_column_f32_scalar, wasm_simd.rs:1691-1784)
// and the quadratic bias inversion (compute.rs:947-998).
//
// MI355X form:
//  * Individuals are put in bucket order (stable within a bucket) within
//    chunks of kSkMembers members.  A bucket's members are then one
//    contiguous run of each chunk's positions.
//  * The signed bucket sums are a product with a sparse +-1 matrix P.  For
//    each block of 32 buckets only the K-steps covering its members are
//    visited; P's non-zero 32-bucket tiles are precomputed as MFMA
//    A-operands, and the MFMA gives exact integers
//      SG[b, j] = sum_{i in b} s_i g_ij,  SM[b, j] = sum_{i in b} s_i [present].
//    For d <= 256 (kSfFusedMaxBlocks blocks) with chunk-local order this is
//    one pass over the .bed windows (k_sketch_fused: LDS-staged rows, gather
//    by precomputed word codes, int8 MFMA on int8 tiles of 32 positions);
//    otherwise the rows are first tiled in bucket order as dosage values
//    (k_gather_chunk, or k_gather4 with --keep) and k_sketch reads the tiles
//    on the e2m1 MFMA (e2m1 tiles of 64 positions; SG = 2 PX - 6 PQ and
//    SM = T_b - 2 PQ from the X = v/2 and Q = [v = 3]/2 operands of the exact
//    band kernel).  For d <= 224 the one pass skips the bucket order
//    altogether (k_sketch_dense: natural order, the +-1 bucket matrix built
//    in LDS, the e2m1 MFMA over every block).
//  * The bucket value y_b = inv_j (SG - mean_j SM) (the reference's f32 mean
//    and inverse std: the same real number its f32 LUT scatter accumulates)
//    is kept as two int8 integers, w = SG - k SM and e = T_b - SM with
//    k = rint(mean) (kSketchI8); a value outside int8 switches the call to
//    the f32 form (kSketchF32: f32 bucket values, bf16 three-part split
//    contraction k_band_sk).  The renormalization to ||y||^2 = N is a
//    per-SNP scale sc_j.
//  * The d-length contraction runs over 64x64 SNP tiles of the band as four
//    exact int8 products per pair (k_band_sk8_2w), finished in FP64; each
//    dot product is scaled by sc_i sc_j, then the quadratic inversion, the
//    unbiased r^2, masking and the annotation contraction shared with the
//    exact path follow.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "band_common.h"

namespace ldsc {

void SketchPlan::release() {
    if (d_fam) (void)hipFree(d_fam);
    if (d_local) (void)hipFree(d_local);
    if (d_ptiles) (void)hipFree(d_ptiles);
    if (d_ptiles4) (void)hipFree(d_ptiles4);
    d_ptiles4 = nullptr;
    if (d_blocks) (void)hipFree(d_blocks);
    if (d_tile_ks) (void)hipFree(d_tile_ks);
    if (d_tsum) (void)hipFree(d_tsum);
    if (d_bo) (void)hipFree(d_bo);
    if (d_bstart) (void)hipFree(d_bstart);
    if (d_pcode) (void)hipFree(d_pcode);
    d_pcode = nullptr;
    if (d_miss) (void)hipFree(d_miss);
    d_miss = nullptr;
    d_tsum = nullptr;
    d_bo = nullptr;
    d_bstart = nullptr;
    d_fam = nullptr;
    d_local = nullptr;
    d_ptiles = nullptr;
    d_blocks = nullptr;
    d_tile_ks = nullptr;
}

// Builds (once per panel and d) the bucket order of individuals and the P
// tiles (int8 for k_sketch_fused, e2m1 for k_sketch).  Host work is
// O(N + chunks * d) plus 2 KiB of tile entries per tile.
int sketch_plan(const ldsc_panel* panel_c, uint64_t d, SketchPlan** out, char* err,
                uint64_t err_len) {
    ldsc_panel* panel = const_cast<ldsc_panel*>(panel_c);
    std::lock_guard<std::mutex> lk(panel->sketch_mu);
    for (SketchPlan* p : panel->sketch_plans)
        if (p->d == d) {
            *out = p;
            return LDSC_OK;
        }
    const uint64_t n = panel->n;
    std::vector<uint32_t> bucket(n);
    std::vector<int8_t> sign(n);
    countsketch_hash(n, d, 42, bucket.data(), sign.data());  // CountSketchProj::new(n, d, 42)
    auto* sp = new SketchPlan();
    sp->d = d;
    sp->d_pad = round_up(d, 32);
    sp->nblocks = sp->d_pad / 32;
    const bool local = panel->keep.empty() && n > kSkMembers;
    sp->chunk = local ? kSkMembers : round_up(n, kKBlock);
    sp->nchunks = (n + sp->chunk - 1) / sp->chunk;
    // positions c*chunk .. : the chunk's members, stable-sorted by bucket;
    // bo[c][B] = first position of bucket block B in chunk c
    std::vector<uint32_t> perm(n);
    std::vector<std::vector<uint64_t>> bo(sp->nchunks, std::vector<uint64_t>(sp->nblocks + 1));
    {
        std::vector<uint64_t> off(d + 1);
        for (uint64_t c = 0; c < sp->nchunks; ++c) {
            const uint64_t c0 = c * sp->chunk, c1 = std::min<uint64_t>(n, c0 + sp->chunk);
            std::fill(off.begin(), off.end(), 0);
            for (uint64_t i = c0; i < c1; ++i) off[bucket[i] + 1]++;
            off[0] = c0;
            for (uint64_t b = 0; b < d; ++b) off[b + 1] += off[b];
            for (uint64_t B = 0; B <= sp->nblocks; ++B) bo[c][B] = off[std::min<uint64_t>(32 * B, d)];
            for (uint64_t i = c0; i < c1; ++i) perm[off[bucket[i]]++] = (uint32_t)i;
        }
    }
    // T_b = sum of the signs of bucket b's members, and sum_b T_b^2
    std::vector<int32_t> tsum(sp->d_pad, 0);
    for (uint64_t i = 0; i < n; ++i) tsum[bucket[i]] += sign[i];
    for (uint64_t b = 0; b < d; ++b) sp->tt += (double)tsum[b] * (double)tsum[b];
    // P tiles, in (bucket block, chunk, K-step) order
    std::vector<int4> blocks(sp->nblocks);
    std::vector<int32_t> tile_ks;
    std::vector<int32_t> bstart(sp->nblocks * sp->nchunks);
    for (uint64_t B = 0; B < sp->nblocks; ++B) {
        blocks[B] = make_int4((int)tile_ks.size(), 0, 0, 0);
        for (uint64_t c = 0; c < sp->nchunks; ++c) {
            const uint64_t lo = bo[c][B], hi = bo[c][B + 1];
            bstart[B * sp->nchunks + c] = (int32_t)tile_ks.size();
            if (hi > lo)
                for (uint64_t ks = lo / 32; ks < (hi + 31) / 32; ++ks) tile_ks.push_back((int32_t)ks);
        }
        blocks[B].y = (int)tile_ks.size() - blocks[B].x;
    }
    std::vector<uint32_t> bo_local(sp->nchunks * (sp->nblocks + 1));
    for (uint64_t c = 0; c < sp->nchunks; ++c)
        for (uint64_t B = 0; B <= sp->nblocks; ++B)
            bo_local[c * (sp->nblocks + 1) + B] = (uint32_t)(bo[c][B] - c * sp->chunk);
    sp->ntiles = tile_ks.size();
    // A-operand tile of (bucket block B, K-step ks): lane l is bucket row
    // 32B + l%32; its byte p is position ks*32 + (l/32)*16 + p.
    std::vector<int8_t> tiles(std::max<uint64_t>(sp->ntiles, 1) * 1024, 0);
    for (uint64_t B = 0; B < sp->nblocks; ++B)
        for (int t = blocks[B].x; t < blocks[B].x + blocks[B].y; ++t) {
            int8_t* tp = tiles.data() + (size_t)t * 1024;
            const uint64_t ks = (uint64_t)tile_ks[t];
            for (int lane = 0; lane < 64; ++lane) {
                const uint64_t b = 32 * B + (lane & 31);
                for (int p = 0; p < 16; ++p) {
                    const uint64_t k = ks * 32 + (lane >> 5) * 16 + p;
                    if (k < n && bucket[perm[k]] == b) tp[lane * 16 + p] = sign[perm[k]];
                }
            }
        }
    // e2m1 A-operand tiles of k_sketch, in (bucket block, chunk, K-step of
    // 64 positions) order: lane l is bucket row 32B + l%32, and nibble jj of
    // its dword q is position ks*64 + (l/32)*32 + 16(q/2) + 2jj + q%2 (the
    // operand placement of the dosage-value tiles), s_i as e2m1 (0x2 = +1,
    // 0xA = -1).
    std::vector<int4> blocks4(sp->nblocks);
    std::vector<int32_t> tile_ks4;
    for (uint64_t B = 0; B < sp->nblocks; ++B) {
        blocks4[B] = make_int4((int)tile_ks4.size(), 0, 0, 0);
        for (uint64_t c = 0; c < sp->nchunks; ++c) {
            const uint64_t lo = bo[c][B], hi = bo[c][B + 1];
            if (hi > lo)
                for (uint64_t ks = lo / 64; ks < (hi + 63) / 64; ++ks) tile_ks4.push_back((int32_t)ks);
        }
        blocks4[B].y = (int)tile_ks4.size() - blocks4[B].x;
    }
    std::vector<uint32_t> tiles4(std::max<size_t>(tile_ks4.size(), 1) * 256, 0u);
    for (uint64_t B = 0; B < sp->nblocks; ++B)
        for (int t = blocks4[B].x; t < blocks4[B].x + blocks4[B].y; ++t) {
            uint32_t* tp = tiles4.data() + (size_t)t * 256;
            const uint64_t ks = (uint64_t)tile_ks4[t];
            for (int lane = 0; lane < 64; ++lane) {
                const uint64_t b = 32 * B + (lane & 31);
                for (int q = 0; q < 4; ++q)
                    for (int jj = 0; jj < 8; ++jj) {
                        const uint64_t k = ks * 64 + (lane >> 5) * 32 + 16 * (q >> 1) + 2 * jj + (q & 1);
                        if (k < n && bucket[perm[k]] == b)
                            tp[lane * 4 + q] |= (sign[perm[k]] < 0 ? 0xAu : 0x2u) << (4 * jj);
                    }
            }
        }
    if (tile_ks4.empty()) tile_ks4.push_back(0);
    // gather maps
    std::vector<int32_t> fam;
    std::vector<uint16_t> loc;
    if (local) {
        loc.assign(round_up(n, kKBlock), kGcPastCode);
        for (uint64_t k = 0; k < n; ++k)
            loc[k] = gc_code((uint32_t)(perm[k] - (k / sp->chunk) * sp->chunk));
    } else {
        std::vector<int32_t> fam_rows(n);
        for (uint64_t k = 0; k < n; ++k)
            fam_rows[k] = panel->keep.empty() ? (int32_t)perm[k] : (int32_t)panel->keep[perm[k]];
        fam = lane_major_map(fam_rows);
    }
    hipError_t e = hipSetDevice(panel->device);
    if (local) {
        if (e == hipSuccess) e = hipMalloc(&sp->d_local, loc.size() * 2);
        if (e == hipSuccess)
            e = hipMemcpy(sp->d_local, loc.data(), loc.size() * 2, hipMemcpyHostToDevice);
    } else {
        if (e == hipSuccess) e = hipMalloc(&sp->d_fam, fam.size() * 4);
        if (e == hipSuccess)
            e = hipMemcpy(sp->d_fam, fam.data(), fam.size() * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc(&sp->d_tile_ks, tile_ks4.size() * 4);
    if (e == hipSuccess)
        e = hipMemcpy(sp->d_tile_ks, tile_ks4.data(), tile_ks4.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&sp->d_ptiles4, tiles4.size() * 4);
    if (e == hipSuccess)
        e = hipMemcpy(sp->d_ptiles4, tiles4.data(), tiles4.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&sp->d_ptiles, tiles.size());
    if (e == hipSuccess) e = hipMemcpy(sp->d_ptiles, tiles.data(), tiles.size(), hipMemcpyHostToDevice);
    if (local) {
        if (e == hipSuccess) e = hipMalloc(&sp->d_bo, bo_local.size() * 4);
        if (e == hipSuccess)
            e = hipMemcpy(sp->d_bo, bo_local.data(), bo_local.size() * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&sp->d_bstart, bstart.size() * 4);
        if (e == hipSuccess)
            e = hipMemcpy(sp->d_bstart, bstart.data(), bstart.size() * 4, hipMemcpyHostToDevice);
    }
    if (local && sp->nblocks <= kSdMaxBlocks) {  // the one-pass form's condition
        // entry of individual i in the P operand of its pair of K-steps
        // (k_sketch_dense): K-step 2(i / 128) + kk, lane half hh, lane
        // bucket row b % 32; within the lane's four words the nibble order of
        // the B operand (word q, nibble jj <-> individual o of the lane's 32)
        std::vector<uint16_t> pc(n);
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t off = (uint32_t)(i % 128), hh = off / 64, kk = (off % 64) / 32;
            const uint32_t o = off % 32, b = bucket[i], B = b / 32;
            const uint32_t q = (o < 16 ? 0u : 2u) + (o & 1u), jj = (o % 16) / 2;
            const uint32_t lane = b % 32 + 32 * hh;
            pc[i] = (uint16_t)(lane | (q << 6) | (kk << 8) | (B << 9) | (jj << 12) |
                               ((sign[i] < 0 ? 1u : 0u) << 15));
        }
        if (e == hipSuccess) e = hipMalloc(&sp->d_pcode, pc.size() * 2);
        if (e == hipSuccess)
            e = hipMemcpy(sp->d_pcode, pc.data(), pc.size() * 2, hipMemcpyHostToDevice);
        const size_t miss = (n + 127) / 128 * 32;
        if (e == hipSuccess) e = hipMalloc(&sp->d_miss, miss);
        if (e == hipSuccess) e = hipMemset(sp->d_miss, 0x55, miss);
    }
    if (e == hipSuccess) e = hipMalloc(&sp->d_tsum, tsum.size() * 4);
    if (e == hipSuccess)
        e = hipMemcpy(sp->d_tsum, tsum.data(), tsum.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&sp->d_blocks, blocks4.size() * sizeof(int4));
    if (e == hipSuccess)
        e = hipMemcpy(sp->d_blocks, blocks4.data(), blocks4.size() * sizeof(int4),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        sp->release();
        delete sp;
        return fail(err, err_len, e == hipErrorOutOfMemory ? LDSC_ERR_OOM : LDSC_ERR_DEVICE,
                    "sketch plan upload failed: %s", hipGetErrorString(e));
    }
    panel->sketch_plans.push_back(sp);
    *out = sp;
    return LDSC_OK;
}

// K-steps whose operands one k_sketch wave has in flight together.
constexpr int kSkAhead = 8;

static inline uint32_t sketch_chunks(const SketchPlan* sp) {
    return (uint32_t)((sp->nblocks + kSkChunk - 1) / kSkChunk);
}

// Share slots of the norm buffer's layout (the stride of its regions).
static inline uint32_t sketch_slots(const SketchPlan* sp) {
    return sketch_share_slots(sp->nblocks);
}

// Layout of the norm buffer: the norm shares (slots x m_pad doubles), then
// for kSketchF32 the scales (m_pad floats), for kSketchI8 the shares of A
// (slots x m_pad doubles), the per-SNP constants (3 x m_pad doubles) and
// the range flag.  Each writer fills its own number of slots from the
// front.
size_t sketch_norm_bytes(const SketchPlan* sp, uint64_t m_pad) {
    const size_t shares = (size_t)sketch_slots(sp) * m_pad * 8;
    return shares + std::max<size_t>(m_pad * 4, shares + 3 * m_pad * 8 + 16);
}

// Wave (gp, c) handles the SNP groups 2gp, 2gp+1 (64 SNPs) and bucket
// blocks [c * kSkChunk, (c+1) * kSkChunk).  Within a block it walks the
// block's e2m1 P tiles (each with its K-step of 64 bucket-order positions)
// with the next kSkAhead tiles' loads in flight, and multiplies each with
// the X and Q operands of the two groups' dosage-value tiles.  The exact
// integer sums SG, SM of a block are then written as:
//  * kSketchRaw: the integer bucket sums (d x m column-major);
//  * kSketchF32: unscaled f32 bucket values y = inv (SG - mean SM) into the
//    sketch tiles sk[((g * nblocks + B) * 64 + lane) * 16 + r] (bucket
//    32B + 8(r/4) + 4(lane/32) + r%4 of SNP 32g + lane%32), and the wave's
//    share of each SNP's squared norm into norm[c * m_pad + snp];
//  * kSketchI8: the same value as two small integers.  With k = rint(mean)
//    and delta = mean - k, and T_b the sign sum of bucket b,
//      w = SG - k SM,  e = T_b - SM  (the signs of b's missing members),
//      y / inv = w + delta e - delta T_b,
//    w and e go to int8 tiles of the same lane order (W at `cols`, E m_pad x
//    d_pad bytes after it), the wave's shares of the squared norm and of
//    A = sum_b (w + delta e) T_b to norm[] and asum[], and a value of a real
//    SNP outside [-127, 127] sets *flag.
//  k_sketch_scale then forms each SNP's renorm scale (and constants).
template <int MODE>
__global__ __launch_bounds__(64) void k_sketch(const uint32_t* __restrict__ tiled, uint32_t nkb,
                                                const v4i* __restrict__ ptiles,
                                                const int4* __restrict__ blocks,
                                                const int32_t* __restrict__ tile_ks,
                                                uint32_t nblocks,
                                                const float2* __restrict__ musc, uint32_t m,
                                                uint32_t m_pad, uint32_t d,
                                                double* __restrict__ norm, void* __restrict__ cols,
                                                int32_t* __restrict__ sg_out,
                                                int32_t* __restrict__ sm_out,
                                                const int32_t* __restrict__ tsum,
                                                double* __restrict__ asum,
                                                int32_t* __restrict__ flag) {
    const uint32_t gp = blockIdx.x, c = blockIdx.y;
    const int lane = threadIdx.x;
    const int h = lane >> 5;
    // the lane's dosage-value dwords of K-block 0 of groups 2gp, 2gp + 1
    const uint32_t* v0p = tiled + ((size_t)(2 * gp) * nkb * 64 + lane) * 4;
    const uint32_t* v1p = v0p + (size_t)nkb * 256;
    const uint32_t snp0 = 64 * gp + (lane & 31), snp1 = snp0 + 32;
    const float2 ms0 = musc[snp0], ms1 = musc[snp1];
    double nrm0 = 0.0, nrm1 = 0.0, as0 = 0.0, as1 = 0.0;
    // integer form: k = rint(mean) in {0, 1, 2} and delta = mean - k (exact)
    const float kf0 = rintf(ms0.x), kf1 = rintf(ms1.x);
    const int k0 = (int)kf0, k1 = (int)kf1;
    const double dl0 = (double)ms0.x - (double)kf0, dl1 = (double)ms1.x - (double)kf1;
    bool bad = false;
    const uint32_t B1 = min(nblocks, (c + 1) * kSkChunk);
    for (uint32_t B = c * kSkChunk; B < B1; ++B) {
        const int4 blk = blocks[B];
        v16f px0, pq0, px1, pq1;
#pragma unroll
        for (int r = 0; r < 16; ++r) px0[r] = pq0[r] = px1[r] = pq1[r] = 0.0f;
        // K-steps (64 positions) in groups of kSkAhead, the next group's
        // operands in flight; steps past the block's last get P = 0 (they add
        // nothing)
        uint2 w0[kSkAhead], w1[kSkAhead];
        v4i P[kSkAhead];
        auto fetch = [&](int s0) {
#pragma unroll
            for (int u = 0; u < kSkAhead; ++u) {
                if (s0 + u < blk.y) {
                    const uint32_t ks = (uint32_t)tile_ks[blk.x + s0 + u];
                    const size_t off = (size_t)(ks >> 1) * 256 + 2 * (ks & 1);
                    w0[u] = *reinterpret_cast<const uint2*>(v0p + off);
                    w1[u] = *reinterpret_cast<const uint2*>(v1p + off);
                    P[u] = ptiles[(size_t)(blk.x + s0 + u) * 64 + lane];
                } else {
                    w0[u] = w1[u] = make_uint2(0u, 0u);
                    P[u] = v4i{0, 0, 0, 0};
                }
            }
        };
        fetch(0);
        for (int s0 = 0; s0 < blk.y; s0 += kSkAhead) {
            uint2 cw0[kSkAhead], cw1[kSkAhead];
            v4i cP[kSkAhead];
#pragma unroll
            for (int u = 0; u < kSkAhead; ++u) {
                cw0[u] = w0[u];
                cw1[u] = w1[u];
                cP[u] = P[u];
            }
            if (s0 + kSkAhead < blk.y) fetch(s0 + kSkAhead);
#pragma unroll
            for (int u = 0; u < kSkAhead; ++u) {
                v8i X0, Q0, X1, Q1;
                vcode_operands(cw0[u].x, cw0[u].y, X0, Q0);
                vcode_operands(cw1[u].x, cw1[u].y, X1, Q1);
                const v8i a8 = __builtin_shufflevector(cP[u], cP[u], 0, 1, 2, 3, -1, -1, -1, -1);
                px0 = LDSC_MFMA_FP4(a8, X0, px0);
                pq0 = LDSC_MFMA_FP4(a8, Q0, pq0);
                px1 = LDSC_MFMA_FP4(a8, X1, px1);
                pq1 = LDSC_MFMA_FP4(a8, Q1, pq1);
            }
        }
        // T_b of the lane's buckets 32B + 8q + 4h + t (t = 0..3), and the
        // exact sums: with PX = sum s v/2 and PQ = sum s [missing]/2 (f32,
        // multiples of 1/2 below 1.5 N), SG = 2 PX - 6 PQ, SM = T_b - 2 PQ
        int T[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 t4 = *reinterpret_cast<const int4*>(tsum + 32 * B + 8 * q + 4 * h);
            T[4 * q] = t4.x;
            T[4 * q + 1] = t4.y;
            T[4 * q + 2] = t4.z;
            T[4 * q + 3] = t4.w;
        }
        v16i sg0, sm0, sg1, sm1;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int q0 = (int)(2.0f * pq0[r]), q1 = (int)(2.0f * pq1[r]);
            sg0[r] = (int)(2.0f * px0[r]) - 3 * q0;
            sm0[r] = T[r] - q0;
            sg1[r] = (int)(2.0f * px1[r]) - 3 * q1;
            sm1[r] = T[r] - q1;
        }
        if (MODE == kSketchRaw) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t b = 32 * B + 8 * (r >> 2) + 4 * h + (r & 3);
                if (b < d) {
                    if (snp0 < m) {
                        sg_out[(size_t)snp0 * d + b] = sg0[r];
                        sm_out[(size_t)snp0 * d + b] = sm0[r];
                    }
                    if (snp1 < m) {
                        sg_out[(size_t)snp1 * d + b] = sg1[r];
                        sm_out[(size_t)snp1 * d + b] = sm1[r];
                    }
                }
            }
        } else if (MODE == kSketchF32) {
            float* sk = static_cast<float*>(cols);
            float* t0 = sk + (((size_t)(2 * gp) * nblocks + B) * 64 + lane) * 16;
            float* t1 = sk + (((size_t)(2 * gp + 1) * nblocks + B) * 64 + lane) * 16;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float y0 = (float)((double)ms0.y * ((double)sg0[r] - (double)ms0.x * (double)sm0[r]));
                const float y1 = (float)((double)ms1.y * ((double)sg1[r] - (double)ms1.x * (double)sm1[r]));
                nrm0 += (double)y0 * (double)y0;
                nrm1 += (double)y1 * (double)y1;
                t0[r] = y0;
                t1[r] = y1;
            }
        } else {
            uint32_t wd0[4] = {0u, 0u, 0u, 0u}, ed0[4] = {0u, 0u, 0u, 0u};
            uint32_t wd1[4] = {0u, 0u, 0u, 0u}, ed1[4] = {0u, 0u, 0u, 0u};
            uint32_t out0 = 0u, out1 = 0u;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int w0 = sg0[r] - k0 * sm0[r], e0 = T[r] - sm0[r];
                const int w1 = sg1[r] - k1 * sm1[r], e1 = T[r] - sm1[r];
                out0 |= ((uint32_t)(w0 + 127) > 254u) | ((uint32_t)(e0 + 127) > 254u);
                out1 |= ((uint32_t)(w1 + 127) > 254u) | ((uint32_t)(e1 + 127) > 254u);
                wd0[r >> 2] |= ((uint32_t)w0 & 0xFFu) << (8 * (r & 3));
                ed0[r >> 2] |= ((uint32_t)e0 & 0xFFu) << (8 * (r & 3));
                wd1[r >> 2] |= ((uint32_t)w1 & 0xFFu) << (8 * (r & 3));
                ed1[r >> 2] |= ((uint32_t)e1 & 0xFFu) << (8 * (r & 3));
                const float y0 = (float)((double)ms0.y * ((double)sg0[r] - (double)ms0.x * (double)sm0[r]));
                const float y1 = (float)((double)ms1.y * ((double)sg1[r] - (double)ms1.x * (double)sm1[r]));
                nrm0 += (double)y0 * (double)y0;
                nrm1 += (double)y1 * (double)y1;
                as0 += ((double)w0 + dl0 * (double)e0) * (double)T[r];
                as1 += ((double)w1 + dl1 * (double)e1) * (double)T[r];
            }
            bad = bad || (out0 && snp0 < m) || (out1 && snp1 < m);
            uint4* W = static_cast<uint4*>(cols);
            uint4* E = W + (size_t)m_pad * nblocks * 2;  // m_pad x d_pad bytes later
            const size_t o0 = ((size_t)(2 * gp) * nblocks + B) * 64 + lane;
            const size_t o1 = ((size_t)(2 * gp + 1) * nblocks + B) * 64 + lane;
            W[o0] = make_uint4(wd0[0], wd0[1], wd0[2], wd0[3]);
            E[o0] = make_uint4(ed0[0], ed0[1], ed0[2], ed0[3]);
            W[o1] = make_uint4(wd1[0], wd1[1], wd1[2], wd1[3]);
            E[o1] = make_uint4(ed1[0], ed1[1], ed1[2], ed1[3]);
        }
    }
    if (MODE != kSketchRaw) {
        nrm0 += __shfl_xor(nrm0, 32);
        nrm1 += __shfl_xor(nrm1, 32);
        if (h == 0) {
            norm[(size_t)c * m_pad + snp0] = nrm0;
            norm[(size_t)c * m_pad + snp1] = nrm1;
        }
    }
    if (MODE == kSketchI8) {
        as0 += __shfl_xor(as0, 32);
        as1 += __shfl_xor(as1, 32);
        if (h == 0) {
            asum[(size_t)c * m_pad + snp0] = as0;
            asum[(size_t)c * m_pad + snp1] = as1;
        }
        if (bad) flag[0] = 1;  // a plain store: every writer stores 1
    }
}

// Ratio-estimator renorm ||col||^2 = N (wasm_simd.rs:1774-1783): the squared
// norm is the chunk shares added in a fixed order, and the column's scale
// sqrt(N / ||y||^2) (f32) is applied by the contraction to each dot
// product (sc_i sc_j <y_i, y_j>) rather than to the stored values.
// kSketchI8 also forms, per SNP, the constants of k_band_sk8: scale = sc
// inv (the renorm scale times the f32 inverse std), delta and A (its shares
// added in the same fixed order).
template <int MODE>
__global__ __launch_bounds__(256) void k_sketch_scale(const double* __restrict__ norm,
                                                       uint32_t nchunks, uint32_t m_pad,
                                                       double n_norm, float* __restrict__ sc,
                                                       const float2* __restrict__ musc,
                                                       const double* __restrict__ asum,
                                                       double* __restrict__ consts) {
    const uint32_t snp = blockIdx.x * blockDim.x + threadIdx.x;
    if (snp >= m_pad) return;
    double nrm = 0.0;
    for (uint32_t c = 0; c < nchunks; ++c) nrm += norm[(size_t)c * m_pad + snp];
    const float s = nrm > 0.0 ? (float)sqrt(n_norm / nrm) : 1.0f;
    if (MODE == kSketchF32) {
        sc[snp] = s;
    } else {
        double A = 0.0;
        for (uint32_t c = 0; c < nchunks; ++c) A += asum[(size_t)c * m_pad + snp];
        const float2 ms = musc[snp];
        consts[snp] = (double)s * (double)ms.y;
        consts[m_pad + snp] = (double)ms.x - (double)rintf(ms.x);
        consts[2 * (size_t)m_pad + snp] = A;
    }
}

__global__ void k_sketch_clear_flag(int32_t* flag) {
    if (threadIdx.x == 0) flag[0] = 0;
}

const float* sketch_scales(const SketchPlan* sp, const double* norm, uint64_t m_pad) {
    return reinterpret_cast<const float*>(norm + (size_t)sketch_slots(sp) * m_pad);
}

const double* sketch_i8_consts(const SketchPlan* sp, const double* norm, uint64_t m_pad) {
    return norm + 2 * (size_t)sketch_slots(sp) * m_pad;
}

int32_t* sketch_i8_flag(const SketchPlan* sp, double* norm, uint64_t m_pad) {
    return reinterpret_cast<int32_t*>(norm + 2 * (size_t)sketch_slots(sp) * m_pad + 3 * m_pad);
}

hipError_t launch_sketch(const SketchPlan* sp, int mode, const uint32_t* tiled, uint32_t nkb,
                         const float2* musc, uint32_t m, uint32_t m_pad, double n, double* norm,
                         void* cols, int32_t* sg, int32_t* sm, hipStream_t st) {
    const uint32_t nch = sketch_chunks(sp), nb = (uint32_t)sp->nblocks;
    const dim3 grid(m_pad / 64, nch), block(64);
    if (mode == kSketchRaw) {
        hipLaunchKernelGGL(k_sketch<kSketchRaw>, grid, block, 0, st, tiled, nkb,
                           (const v4i*)sp->d_ptiles4, sp->d_blocks, sp->d_tile_ks, nb, musc, m, m_pad,
                           (uint32_t)sp->d, nullptr, nullptr, sg, sm, sp->d_tsum, nullptr, nullptr);
        return hipGetLastError();
    }
    hipError_t e;
    if (mode == kSketchF32) {
        hipLaunchKernelGGL(k_sketch<kSketchF32>, grid, block, 0, st, tiled, nkb,
                           (const v4i*)sp->d_ptiles4, sp->d_blocks, sp->d_tile_ks, nb, musc, m, m_pad,
                           (uint32_t)sp->d, norm, cols, nullptr, nullptr, sp->d_tsum, nullptr, nullptr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_sketch_scale<kSketchF32>, dim3((m_pad + 255) / 256), dim3(256), 0, st,
                           norm, nch, m_pad, n, const_cast<float*>(sketch_scales(sp, norm, m_pad)),
                           nullptr, nullptr, nullptr);
        return hipGetLastError();
    }
    double* asum = norm + (size_t)sketch_slots(sp) * m_pad;
    double* consts = const_cast<double*>(sketch_i8_consts(sp, norm, m_pad));
    int32_t* flag = sketch_i8_flag(sp, norm, m_pad);
    hipLaunchKernelGGL(k_sketch_clear_flag, dim3(1), dim3(64), 0, st, flag);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sketch<kSketchI8>, grid, block, 0, st, tiled, nkb,
                       (const v4i*)sp->d_ptiles4, sp->d_blocks, sp->d_tile_ks, nb, musc, m, m_pad,
                       (uint32_t)sp->d, norm, cols, nullptr, nullptr, sp->d_tsum, asum, flag);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sketch_scale<kSketchI8>, dim3((m_pad + 255) / 256), dim3(256), 0, st,
                       norm, nch, m_pad, n, nullptr, musc, asum, consts);
    return hipGetLastError();
}

// ------------------------------------------------------------------------
// One-pass bucket sums (chunk-local bucket order, integer-operand form)
// ------------------------------------------------------------------------
// Workgroup (g, r) owns the 32 SNPs of group g and bucket blocks
// [8 BPW r, 8 BPW (r+1)); its wave v owns BPW consecutive blocks of them.
// For each member chunk c (kSkMembers FAM rows, one 4 KiB window of each
// .bed row) the workgroup
//  1. loads the window of its 32 rows (wave v: rows 4v .. 4v+3, one 16-byte
//     load per lane and row, the next pass's loads in flight), counts their
//     integer stats and stages them in LDS byte-interleaved by row quad
//     (word k of quad Q = byte k of rows 4Q .. 4Q+3, in 16-byte slots placed
//     by gc_slot) -- as k_gather_chunk;
//  2. per wave, walks the K-steps (32 positions of the chunk's bucket order)
//     that its blocks cover, four at a time.  A lane gathers 16 positions
//     (one K-step half) of one row quad: the plan's gc_code of each position
//     gives the member's LDS word and field, and one 4-byte read its code in
//     the quad's four rows.  The row dwords, in the tiles' cell order, are
//     staged; the wave then runs, for each K-step and each of its blocks the
//     K-step reaches, v_mfma_i32_32x32x32_i8 of that block's +-1 tile with
//     the step's G and M operands.  The first block's tiles of the four
//     steps load before the gather.
// The exact SG, SM of every block are complete after the last chunk; the
// epilogue forms w, e, the norm and A shares and the range flag exactly as
// k_sketch<kSketchI8>, with the f32 mean / inverse std from the stats the
// workgroup counted itself (the same operations as k_snp_info).  No tiled
// copy of the rows is written.  Workgroups are mapped XCD-aware so that the
// ranges of one group (which stage the same windows) share an L2.
constexpr uint32_t kSfWords = kSkMembers / 4;   // LDS words per quad and chunk
constexpr uint32_t kSfStride = kSfWords + 4;    // word kSfWords reads as code 01
constexpr uint32_t kSfSteps = 4;                // K-steps per gather pass
constexpr uint32_t kSfStage = kSfSteps * 64;    // staging dwords per wave

// RAW: the epilogue writes the exact sums SG, SM (d x m column-major int32,
// sg_out / sm_out) instead of the int8 form, for ldsc_l2_sketch_buckets.
template <int BPW, bool RAW>
__global__ __launch_bounds__(512) void k_sketch_fused(
    const uint8_t* __restrict__ slabs, uint64_t pitch, uint32_t src_dw, uint32_t n,
    const uint64_t* __restrict__ bed_idx, uint64_t base, uint32_t m, uint32_t m_pad,
    const uint16_t* __restrict__ local, uint32_t nchunks, const v4i* __restrict__ ptiles,
    const uint32_t* __restrict__ bo, const int32_t* __restrict__ bstart, uint32_t nblocks,
    uint32_t nranges, uint32_t ngroups, uint32_t per_xcd, const int32_t* __restrict__ tsum,
    uint32_t* __restrict__ stats_out, double* __restrict__ norm, double* __restrict__ asum,
    void* __restrict__ cols, int32_t* __restrict__ flag, uint32_t d, int32_t* __restrict__ sg_out,
    int32_t* __restrict__ sm_out) {
    extern __shared__ uint32_t lds[];  // 8 quads x kSfStride, then kSfWaves x kSfStage staging
    __shared__ uint32_t srow[32][3];
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const uint32_t logical = xcd * per_xcd + slot;
    if (logical >= ngroups * nranges) return;
    const uint32_t g = logical / nranges, r = logical % nranges;
    const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63, h = ln >> 5;
    constexpr uint32_t cw = kSkMembers / 16;  // window dwords per row
    constexpr int kPasses = (int)(cw / 256);
    uint32_t* stage = lds + 8 * kSfStride + wv * kSfStage;
    if (ln == 0) lds[wv * kSfStride + kSfWords] = 0x55555555u;
    // load role: rows 4wv + q (quad wv); padding rows read as missing
    const uint32_t* rp[4];
    uint32_t pad_or = 0;  // bit q: row 4wv + q is padding
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t j = 32 * g + 4 * wv + q;
        const uint64_t row = j < m ? (bed_idx ? bed_idx[j] : base + j) : 0;
        rp[q] = reinterpret_cast<const uint32_t*>(slabs + row * pitch);
        pad_or |= (j < m ? 0u : 1u) << q;
    }
    const uint32_t B0 = r * (kSfWaves * BPW) + wv * BPW;
    v16i sg[BPW], sm[BPW];
#pragma unroll
    for (int k = 0; k < BPW; ++k)
#pragma unroll
        for (int x = 0; x < 16; ++x) sg[k][x] = sm[k][x] = 0;
    uint32_t cnt[4][3] = {};
    // gather role of the lane: K-step gi of a pass, row quad gQ, half gh
    const uint32_t gi = ln >> 4, gQ = (ln >> 1) & 7, gh = ln & 1;
    for (uint32_t c = 0; c < nchunks; ++c) {
        __syncthreads();  // the previous chunk's gathers are done with the quads
        // 1. stage the chunk's window of the wave's four rows.  Only the last
        //    chunk holds positions past the last member (read as code 01).
        const bool tail = (uint64_t)(c + 1) * kSkMembers > n;
        auto load_pass = [&](int i, uint4 (&v)[4]) {
            const uint64_t q0 = (uint64_t)c * cw + i * 256 + 4 * ln;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                v[q] = q0 < src_dw ? *reinterpret_cast<const uint4*>(rp[q] + q0)
                                   : make_uint4(0u, 0u, 0u, 0u);
        };
        uint4 vn[4];
        load_pass(0, vn);
#pragma unroll 1
        for (int i = 0; i < kPasses; ++i) {
            const uint32_t k0 = i * 256 + 4 * ln;  // window dword of the lane's first word
            const uint64_t q0 = (uint64_t)c * cw + k0;
            uint4 v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = vn[q];
            if (i + 1 < kPasses) load_pass(i + 1, vn);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                uint32_t w[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t x = t == 0 ? v[q].x : t == 1 ? v[q].y : t == 2 ? v[q].z : v[q].w;
                    if (tail) x = ts_mask(x, q0 + t, n);
                    w[q] = (pad_or >> q) & 1u ? 0x55555555u : x;
                    const uint32_t L = w[q] & 0x55555555u, H = (w[q] >> 1) & 0x55555555u;
                    cnt[q][0] += __popc(~(L | H) & 0x55555555u);  // 00: dosage 2
                    cnt[q][1] += __popc(~L & H);                   // 10: dosage 1
                    cnt[q][2] += __popc(L & H);                    // 11: dosage 0
                }
                uint32_t o[4];
                transpose_bytes4(w[0], w[1], w[2], w[3], o);
                reinterpret_cast<uint4*>(lds + wv * kSfStride)[gc_slot(k0 + t)] =
                    make_uint4(o[0], o[1], o[2], o[3]);
            }
        }
        __syncthreads();
        // 2. the wave's K-steps in this chunk: the union of its blocks' runs
        uint32_t blo[BPW], bhi[BPW];
        int32_t bts[BPW];
        uint32_t ks_lo = 0xFFFFFFFFu, ks_hi = 0;
#pragma unroll
        for (int k = 0; k < BPW; ++k) {
            const uint32_t B = B0 + k;
            blo[k] = bhi[k] = 0;
            bts[k] = 0;
            if (B < nblocks) {
                const uint32_t p0 = bo[(size_t)c * (nblocks + 1) + B];
                const uint32_t p1 = bo[(size_t)c * (nblocks + 1) + B + 1];
                if (p1 > p0) {
                    blo[k] = p0 / 32;
                    bhi[k] = (p1 + 31) / 32;
                    bts[k] = bstart[(size_t)B * nchunks + c];
                    ks_lo = min(ks_lo, blo[k]);
                    ks_hi = max(ks_hi, bhi[k]);
                }
            }
        }
        const uint16_t* lc = local + (size_t)c * kSkMembers;
        uint4 le0 = make_uint4(0u, 0u, 0u, 0u), le1 = le0;  // the pass's 16 codes
        auto load_codes = [&](uint32_t ks0) {
            const uint32_t ks = ks0 + gi;
            if (ks < ks_hi) {
                const uint4* lp = reinterpret_cast<const uint4*>(lc + 32 * ks + 16 * gh);
                le0 = lp[0];
                le1 = lp[1];
            }
        };
        if (ks_lo < ks_hi) load_codes(ks_lo);
        for (uint32_t ks0 = ks_lo; ks0 < ks_hi; ks0 += kSfSteps) {
            // the +-1 tiles of the pass's K-steps for the first of the wave's
            // blocks reaching each, loaded before the gather (further blocks
            // reaching a K-step -- at block boundaries, or several short runs
            // in the last, partial chunk -- load their tiles where used)
            v4i P1[kSfSteps];
            int k1[kSfSteps];
#pragma unroll
            for (int i = 0; i < (int)kSfSteps; ++i) {
                const uint32_t ks = ks0 + i;
                k1[i] = -1;
#pragma unroll
                for (int k = BPW - 1; k >= 0; --k)
                    if (ks >= blo[k] && ks < bhi[k]) k1[i] = k;
                if (k1[i] >= 0)
                    P1[i] = ptiles[(size_t)(bts[k1[i]] + (int32_t)(ks - blo[k1[i]])) * 64 + ln];
            }
            const uint4 e0 = le0, e1 = le1;
            if (ks0 + kSfSteps < ks_hi) load_codes(ks0 + kSfSteps);
            // gather: lane -> (K-step ks0 + gi, row quad gQ, half gh)
            if (ks0 + gi < ks_hi) {
                const uint32_t ev[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
                const uint32_t* quad = lds + gQ * kSfStride;
                uint32_t acc[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int e = 0; e < 16; ++e) {  // gc_code: LDS word * 4 + field
                    const uint32_t code = (ev[e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
                    acc[e & 3] |= ((quad[code >> 2] >> (2 * (code & 3))) & 0x03030303u)
                                  << (2 * (e >> 2));
                }
                uint32_t o[4];
                transpose_bytes4(acc[0], acc[1], acc[2], acc[3], o);  // o[x] = row 4gQ + x
                reinterpret_cast<uint4*>(stage)[(gi * 2 + gh) * 8 + gQ] =
                    make_uint4(o[0], o[1], o[2], o[3]);
            }
            // the staging writes above are read by other lanes of this wave
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int i = 0; i < (int)kSfSteps; ++i) {
                if (k1[i] < 0) continue;  // past the last K-step
                const uint32_t dw = stage[(i * 2 + h) * 32 + (ln & 31)];
                v4i G, M;
                expand_codes(dw, G, M);
                const uint32_t ks = ks0 + i;
#pragma unroll
                for (int k = 0; k < BPW; ++k) {
                    if (k == k1[i]) {
                        sg[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(P1[i], G, sg[k], 0, 0, 0);
                        sm[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(P1[i], M, sm[k], 0, 0, 0);
                    } else if (ks >= blo[k] && ks < bhi[k]) {
                        const v4i P2 = ptiles[(size_t)(bts[k] + (int32_t)(ks - blo[k])) * 64 + ln];
                        sg[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(P2, G, sg[k], 0, 0, 0);
                        sm[k] = __builtin_amdgcn_mfma_i32_32x32x32_i8(P2, M, sm[k], 0, 0, 0);
                    }
                }
            }
            // the next gather rewrites the staging only after these reads
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    const uint32_t row = ln & 31, snp = 32 * g + row;
    if (RAW) {
#pragma unroll
        for (int k = 0; k < BPW; ++k) {
            const uint32_t B = B0 + k;
            if (B >= nblocks || snp >= m) continue;
#pragma unroll
            for (int x = 0; x < 16; ++x) {
                const uint32_t b = 32 * B + 8 * (x >> 2) + 4 * h + (x & 3);
                if (b < d) {
                    sg_out[(size_t)snp * d + b] = sg[k][x];
                    sm_out[(size_t)snp * d + b] = sm[k][x];
                }
            }
        }
        return;
    }
    // stats of the 32 rows over every chunk: wave wv holds rows 4wv .. 4wv+3
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uint32_t x = cnt[q][k];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
            cnt[q][k] = x;
        }
    if (ln < 4) {
        uint32_t n2 = 0, n1 = 0, n0 = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q == ln) {
                n2 = cnt[q][0];
                n1 = cnt[q][1];
                n0 = cnt[q][2];
            }
        srow[4 * wv + ln][0] = n2;
        srow[4 * wv + ln][1] = n1;
        srow[4 * wv + ln][2] = n0;
    }
    __syncthreads();
    const uint32_t n2 = srow[row][0], n1 = srow[row][1], n0 = srow[row][2];
    const uint32_t S = 2 * n2 + n1, C = n2 + n1 + n0, SS = 4 * n2 + n1;
    if (r == 0 && wv == 0 && ln < 32)
        reinterpret_cast<uint4*>(stats_out)[snp] = make_uint4(S, C, SS, n1);
    float a32, i32;
    (void)ref_mean_inv_f32(S, C, SS, n, a32, i32);
    const float kf = rintf(a32);
    const int kk = (int)kf;
    const double dl = (double)a32 - (double)kf;
    double nrm = 0.0, as = 0.0;
    bool bad = false;
    uint4* W = static_cast<uint4*>(cols);
    uint4* E = W + (size_t)m_pad * nblocks * 2;  // m_pad x d_pad bytes later
#pragma unroll
    for (int k = 0; k < BPW; ++k) {
        const uint32_t B = B0 + k;
        if (B >= nblocks) continue;
        int T[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 t4 = *reinterpret_cast<const int4*>(tsum + 32 * B + 8 * q + 4 * h);
            T[4 * q] = t4.x;
            T[4 * q + 1] = t4.y;
            T[4 * q + 2] = t4.z;
            T[4 * q + 3] = t4.w;
        }
        uint32_t wd[4] = {0u, 0u, 0u, 0u}, ed[4] = {0u, 0u, 0u, 0u};
        uint32_t out = 0u;
#pragma unroll
        for (int x = 0; x < 16; ++x) {
            const int w = sg[k][x] - kk * sm[k][x], e = T[x] - sm[k][x];
            out |= ((uint32_t)(w + 127) > 254u) | ((uint32_t)(e + 127) > 254u);
            wd[x >> 2] |= ((uint32_t)w & 0xFFu) << (8 * (x & 3));
            ed[x >> 2] |= ((uint32_t)e & 0xFFu) << (8 * (x & 3));
            const float y = (float)((double)i32 * ((double)sg[k][x] - (double)a32 * (double)sm[k][x]));
            nrm += (double)y * (double)y;
            as += ((double)w + dl * (double)e) * (double)T[x];
        }
        bad = bad || (out && snp < m);
        const size_t o = ((size_t)g * nblocks + B) * 64 + ln;
        W[o] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        E[o] = make_uint4(ed[0], ed[1], ed[2], ed[3]);
    }
    nrm += __shfl_xor(nrm, 32);
    as += __shfl_xor(as, 32);
    if (h == 0) {
        const size_t sl = (size_t)(r * kSfWaves + wv) * m_pad + snp;
        norm[sl] = nrm;
        asum[sl] = as;
    }
    if (bad) flag[0] = 1;  // a plain store: every writer stores 1
}

// ------------------------------------------------------------------------
// Dense one-pass bucket sums (at most kSdMaxBlocks blocks, no --keep)
// ------------------------------------------------------------------------
// The signed bucket sums of a 32-SNP group are the product P . G of the
// group's .bed rows (K = individuals in FAM order) with the bucket matrix
// (P[b][i] = s_i when individual i is in bucket b), on the e2m1 MFMA
// v_mfma_scale_f32_32x32x64_f8f6f4: no gather and no bucket order.  Every
// K-step of 64 individuals meets every bucket block, so the MFMA work is
// two products per block and K-step; it pays while the blocks are few
// (measured: faster than k_sketch_fused up to 7 blocks, slower at 8).
//  * Workgroup q: 8 waves, two on each SIMD.  Waves v and v + 4 own SNP
//    group 4q + v; wave v the first ceil(NB/2) blocks, wave v + 4 the rest
//    (so that no wave holds more than 4 x 2 x 16 accumulators).  The second
//    half also builds P and counts the integer stats, which it passes to
//    the first through LDS at the end.
//  * P for a phase of two pairs of K-steps (256 individuals) is built in
//    LDS from the plan's per-individual code (lane, word, nibble, block and
//    sign of the individual's entry in the A-operand layout), one LDS atomic
//    OR per individual; three buffers rotate (read / being built / being
//    cleared) with one barrier per phase, and a buffer is cleared by XOR-ing
//    out the entries its builders placed two phases before.
//  * The B operands come straight from the .bed rows: lane l holds SNP
//    32g + l%32 and, per pair p, 16 bytes of its row (individuals
//    128p + 64h .. +63, h = l/32: the first 8 bytes K-step 2p, the next 8
//    K-step 2p+1).  Each 2-bit PLINK code c becomes the value code
//    v = c ^ 2 ^ (c >> 1) (the dosage; 3 when missing), split into the nibble
//    operands X = v/2 and Q = [v = 3]/2 as in the exact band kernel.  Rows
//    and codes load two phases ahead; padding SNPs read an all-missing row.
//  * The K loop is software-pipelined by K-step: the next K-step's operands
//    are expanded, and its P tiles read from LDS, while this K-step's MFMAs
//    run.
//  * The f32 sums are exact (multiples of 1/2 below 1.5 N), and
//    SG = 2 PX - 6 PQ, SM = T_b - 2 PQ.  The epilogue is k_sketch_fused's,
//    with one norm share slot per half.
constexpr uint32_t kSdWaves = 4;
constexpr uint32_t kSdPhasePairs = 2;  // pairs of K-steps per barrier interval (phase)

// One .bed dword of the lane's row (16 cells of PLINK codes, L at the even
// and H at the odd bit of each): the nibble operands X (value code v = the
// dosage, 3 when missing; e2m1 v/2) and Q ([v = 3]; e2m1 1/2) of cells 2jj
// (xe, qe) and 2jj + 1 (xo, qo) at nibble jj, and the counters a (codes 01
// and 11), b (codes 10 and 11) and c (code 11).
template <bool STATS>
__device__ __forceinline__ void dense_word(uint32_t w, uint32_t& xe, uint32_t& xo, uint32_t& qe,
                                           uint32_t& qo, uint32_t& a, uint32_t& b, uint32_t& c) {
    const uint32_t t1 = w >> 1;
    const uint32_t hb = t1 & 0x55555555u;         // H, at the cell's even bit
    const uint32_t v = (w ^ hb) ^ 0xAAAAAAAAu;    // (L ^ H) | (~H << 1)
    const uint32_t mq = (w & ~t1) & 0x55555555u;  // L & ~H: code 01
    xe = v & 0x33333333u;
    xo = (v >> 2) & 0x33333333u;
    qe = mq & 0x11111111u;
    qo = (mq >> 2) & 0x11111111u;
    if (STATS) {
        a += __popc(w & 0x55555555u);
        b += __popc(hb);
        c += __popc(w & hb);
    }
}

// The K loop and epilogue of one half of the workgroup: its four waves own
// SNP groups 4 blockIdx.x + (wave & 3) and the NBW blocks B0 .. B0 + NBW - 1.
// The BUILD half also keeps the P buffers (its 256 threads place the 256
// individuals of a phase and clear the entries of two phases before).
template <int NB, int NBW, bool BUILD, bool STATS, bool RAW>
__device__ __forceinline__ void sketch_dense_half(
    uint32_t* lds, const uint8_t* __restrict__ slabs, uint64_t pitch, uint32_t row_dw, uint32_t n,
    const uint64_t* __restrict__ bed_idx, uint64_t base, uint32_t m, uint32_t m_pad,
    const uint16_t* __restrict__ pcode, const uint8_t* __restrict__ miss_row,
    const int32_t* __restrict__ tsum, uint32_t* __restrict__ stats_out,
    double* __restrict__ norm, double* __restrict__ asum, void* __restrict__ cols,
    int32_t* __restrict__ flag, uint32_t d, int32_t* __restrict__ sg_out,
    int32_t* __restrict__ sm_out, uint32_t B0, uint32_t slot) {
    constexpr uint32_t kPair = 2 * NB * 64 * 4;  // P words of a pair of K-steps
    constexpr uint32_t kBuf = kSdPhasePairs * kPair;
    const uint32_t ngroups = m_pad / 32;
    const int tid = threadIdx.x & 255, wv = tid >> 6, ln = tid & 63, h = ln >> 5;
    const uint32_t g = blockIdx.x * kSdWaves + wv;
    const uint32_t snp = 32 * g + (ln & 31);
    const bool real = g < ngroups && snp < m;  // padding SNPs read an all-missing row
    const uint64_t row = real ? (bed_idx ? bed_idx[snp] : base + snp) : 0;
    const uint32_t* rp = reinterpret_cast<const uint32_t*>(real ? slabs + row * pitch : miss_row);
    const uint32_t npairs = (n + 127) / 128;
    const uint32_t nph = (npairs + kSdPhasePairs - 1) / kSdPhasePairs;
    const uint32_t niter = (nph + 1) / 2;  // two phases an iteration
    // Row bytes and codes load unconditionally (from a clamped address) into
    // registers that stay in flight for two phases; what is out of range is
    // replaced where it is used.  The lane's 16 row bytes of pair p are
    // individuals 128p + 64h .. +63.
    auto row_ok = [&](uint32_t p) { return p < npairs && 8 * p + 4 * h + 4 <= row_dw; };
    auto load_row = [&](uint32_t p) {
        return *reinterpret_cast<const uint4*>(rp + (row_ok(p) ? 8 * p + 4 * h : 0));
    };
    // thread t places individual 256 f + t of phase f (code: lane | word << 6 |
    // K-step << 8 | block << 9 | nibble << 12 | negative << 15), in pair t / 128
    auto load_code = [&](uint32_t f) {
        const uint32_t i = 256 * f + tid;
        return (uint32_t)pcode[i < n ? i : 0];
    };
    auto code_ok = [&](uint32_t c, uint32_t f) { return 256 * f + tid < n ? c : 0xFFFFFFFFu; };
    const uint32_t pair_off = (tid >> 7) * kPair;
    auto entry = [&](uint32_t c, uint32_t& word, uint32_t& bits) {
        // an absent code (c = ~0) ORs 0 into the pair's first word
        const bool ok = c != 0xFFFFFFFFu;
        word = pair_off + (ok ? ((((c >> 8) & 1u) * NB + ((c >> 9) & 7u)) * 64 + (c & 63u)) * 4 +
                                    ((c >> 6) & 3u)
                              : 0u);
        bits = ok ? ((c >> 15) & 1u ? 0xAu : 0x2u) << (4 * ((c >> 12) & 7u)) : 0u;
    };
    // Every pair before the last is whole and inside the row (the row
    // stride holds ceil(N / 4) bytes); the last pair's cells past n and the
    // pairs after it read as code 01 (missing).  The branch is uniform.
    auto sanitize = [&](uint4 r, uint32_t p) {
        if (p + 1 >= npairs) {
            const uint32_t q0 = 8 * p + 4 * h;
            if (!row_ok(p)) r = make_uint4(0x55555555u, 0x55555555u, 0x55555555u, 0x55555555u);
            r.x = ts_mask(r.x, q0, n);
            r.y = ts_mask(r.y, q0 + 1, n);
            r.z = ts_mask(r.z, q0 + 2, n);
            r.w = ts_mask(r.w, q0 + 3, n);
        }
        return r;
    };
    constexpr int NBA = NBW > 0 ? NBW : 1;  // array extent
    v16f px[NBA], pq[NBA];
#pragma unroll
    for (int B = 0; B < NBA; ++B)
#pragma unroll
        for (int x = 0; x < 16; ++x) px[B][x] = pq[B][x] = 0.0f;
    uint32_t ca = 0, cb = 0, cc = 0;
    auto expand = [&](uint32_t w0, uint32_t w1, v8i& X, v8i& Q) {
        uint32_t e0, e1, e2, e3, g0, g1, g2, g3;
        dense_word<STATS>(w0, e0, e1, g0, g1, ca, cb, cc);
        dense_word<STATS>(w1, e2, e3, g2, g3, ca, cb, cc);
        const v4i x4 = {(int)e0, (int)e1, (int)e2, (int)e3};
        const v4i q4 = {(int)g0, (int)g1, (int)g2, (int)g3};
        X = __builtin_shufflevector(x4, x4, 0, 1, 2, 3, -1, -1, -1, -1);
        Q = __builtin_shufflevector(q4, q4, 0, 1, 2, 3, -1, -1, -1, -1);
    };
    // P tiles of the wave's blocks for K-step s of the phase in buffer `buf`
    // (pair s / 2, K-step s % 2)
    auto read_p = [&](uint32_t buf, int s, v4i (&a)[NBA]) {
        const uint32_t* src =
            lds + buf + (s >> 1) * kPair + ((s & 1) * NB + B0) * 256 + ln * 4;
#pragma unroll
        for (int B = 0; B < NBW; ++B) a[B] = *reinterpret_cast<const v4i*>(src + B * 256);
    };
    auto mma = [&](const v4i (&a)[NBA], const v8i& X, const v8i& Q) {
#pragma unroll
        for (int B = 0; B < NBW; ++B) {
            const v8i a8 = __builtin_shufflevector(a[B], a[B], 0, 1, 2, 3, -1, -1, -1, -1);
            px[B] = LDSC_MFMA_FP4(a8, X, px[B]);
            pq[B] = LDSC_MFMA_FP4(a8, Q, pq[B]);
        }
    };
    uint4 ring[2 * kSdPhasePairs];  // row bytes, pair 2f + pp in slot 2(f % 2) + pp
    uint32_t cring[2] = {0u, 0u};   // code of phase f + 1, in slot f % 2
    uint32_t cw[2] = {0u, 0u}, cbits[2] = {0u, 0u};  // entry built in phase f - 2 (slot f % 2)
#pragma unroll
    for (int k = 0; k < 2 * (int)kSdPhasePairs; ++k) ring[k] = load_row(k);
    if (BUILD) {
        cring[0] = load_code(1);
        cring[1] = load_code(2);
    }
    for (uint32_t x = 4 * threadIdx.x; x < 3 * kBuf; x += 4 * 64 * 2 * kSdWaves)
        *reinterpret_cast<uint4*>(lds + x) = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
    if (BUILD) {
        uint32_t wd, bt;
        entry(code_ok(load_code(0), 0), wd, bt);
        atomicOr(lds + wd, bt);
        cw[1] = wd;
        cbits[1] = bt;
    }
    __syncthreads();
    uint32_t rd = 0, bd = kBuf, cl = 2 * kBuf;  // read / being built / being cleared
    // Software-pipelined by K-step (four a phase): the next K-step's
    // operands are expanded, and its P tiles read, while this K-step's MFMAs
    // run.  Past the last pair the rows read as all missing and no P entry
    // exists.
    v8i X0, Q0, X1, Q1;
    ring[0] = sanitize(ring[0], 0);
    expand(ring[0].x, ring[0].y, X0, Q0);
    v4i pa[NBA], pb[NBA];
    for (uint32_t it = 0; it < niter; ++it) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint32_t f = 2 * it + u;
            const uint32_t p0 = kSdPhasePairs * f;
            read_p(rd, 0, pa);
            if (BUILD) {
                uint32_t wd, bt;
                entry(code_ok(cring[u], f + 1), wd, bt);
                atomicOr(lds + bd + wd, bt);
                atomicXor(lds + cl + cw[u], cbits[u]);
                cw[u] = wd;
                cbits[u] = bt;
            }
            const uint4 r1 = sanitize(ring[2 * u + 1], p0 + 1);
            read_p(rd, 1, pb);
            expand(ring[2 * u].z, ring[2 * u].w, X1, Q1);
            mma(pa, X0, Q0);
            ring[2 * u] = load_row(p0 + 2 * kSdPhasePairs);
            read_p(rd, 2, pa);
            expand(r1.x, r1.y, X0, Q0);
            mma(pb, X1, Q1);
            read_p(rd, 3, pb);
            expand(r1.z, r1.w, X1, Q1);
            mma(pa, X0, Q0);
            ring[2 * u + 1] = load_row(p0 + 1 + 2 * kSdPhasePairs);
            ring[2 * (1 - u)] = sanitize(ring[2 * (1 - u)], p0 + kSdPhasePairs);
            expand(ring[2 * (1 - u)].x, ring[2 * (1 - u)].y, X0, Q0);
            mma(pb, X1, Q1);
            if (BUILD) cring[u] = load_code(f + 3);
            __syncthreads();
            const uint32_t t = rd;
            rd = bd;
            bd = cl;
            cl = t;
        }
    }
    // The STATS half counted the integer stats (every lane expanded
    // 2 + 16 niter words of 16 cells; the words past the last pair read as
    // missing) and passes them to the other half through LDS.
    uint4 st = make_uint4(0u, 0u, 0u, 0u);
    if (STATS) {
        uint32_t n0 = cc, n1 = cb - cc, n2 = 16 * (2 + 16 * niter) - ca - cb + cc;
        n2 += __shfl_xor(n2, 32);
        n1 += __shfl_xor(n1, 32);
        n0 += __shfl_xor(n0, 32);
        st = make_uint4(2 * n2 + n1, n2 + n1 + n0, 4 * n2 + n1, n1);
        if (h == 0) reinterpret_cast<uint4*>(lds)[wv * 32 + (ln & 31)] = st;
    }
    __syncthreads();
    if (!STATS) st = reinterpret_cast<const uint4*>(lds)[wv * 32 + (ln & 31)];
    if (g >= ngroups) return;
    const uint32_t S = st.x, C = st.y, SS = st.z;
    // T_b of the lane's 16 buckets of block B (bucket 32B + 8(x/4) + 4h + x%4)
    auto load_t = [&](uint32_t B, int (&T)[16]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int4 t4 = *reinterpret_cast<const int4*>(tsum + 32 * B + 8 * q + 4 * h);
            T[4 * q] = t4.x;
            T[4 * q + 1] = t4.y;
            T[4 * q + 2] = t4.z;
            T[4 * q + 3] = t4.w;
        }
    };
    if (RAW) {
        if (snp >= m) return;
#pragma unroll
        for (int B = 0; B < NBW; ++B) {
            int T[16];
            load_t(B0 + B, T);
#pragma unroll
            for (int x = 0; x < 16; ++x) {
                const uint32_t b = 32 * (B0 + B) + 8 * (x >> 2) + 4 * h + (x & 3);
                const int sq = (int)(2.0f * pq[B][x]);
                if (b < d) {
                    sg_out[(size_t)snp * d + b] = (int)(2.0f * px[B][x]) - 3 * sq;
                    sm_out[(size_t)snp * d + b] = T[x] - sq;
                }
            }
        }
        return;
    }
    if (STATS && h == 0) reinterpret_cast<uint4*>(stats_out)[snp] = st;
    float a32, i32;
    (void)ref_mean_inv_f32(S, C, SS, n, a32, i32);
    const float kf = rintf(a32);
    const int kk = (int)kf;
    const double dl = (double)a32 - (double)kf;
    double nrm = 0.0, as = 0.0;
    bool bad = false;
    uint4* W = static_cast<uint4*>(cols);
    uint4* E = W + (size_t)m_pad * NB * 2;  // m_pad x d_pad bytes later
#pragma unroll
    for (int B = 0; B < NBW; ++B) {
        int T[16];
        load_t(B0 + B, T);
        uint32_t wd[4] = {0u, 0u, 0u, 0u}, ed[4] = {0u, 0u, 0u, 0u};
        uint32_t out = 0u;
#pragma unroll
        for (int x = 0; x < 16; ++x) {
            const int sq = (int)(2.0f * pq[B][x]);
            const int sg = (int)(2.0f * px[B][x]) - 3 * sq, sm = T[x] - sq;
            const int w = sg - kk * sm, e = T[x] - sm;
            out |= ((uint32_t)(w + 127) > 254u) | ((uint32_t)(e + 127) > 254u);
            wd[x >> 2] |= ((uint32_t)w & 0xFFu) << (8 * (x & 3));
            ed[x >> 2] |= ((uint32_t)e & 0xFFu) << (8 * (x & 3));
            const float y = (float)((double)i32 * ((double)sg - (double)a32 * (double)sm));
            nrm += (double)y * (double)y;
            as += ((double)w + dl * (double)e) * (double)T[x];
        }
        bad = bad || (out && snp < m);
        const size_t o = ((size_t)g * NB + B0 + B) * 64 + ln;
        W[o] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
        E[o] = make_uint4(ed[0], ed[1], ed[2], ed[3]);
    }
    nrm += __shfl_xor(nrm, 32);
    as += __shfl_xor(as, 32);
    if (h == 0) {
        norm[(size_t)slot * m_pad + snp] = nrm;
        asum[(size_t)slot * m_pad + snp] = as;
    }
    if (bad) flag[0] = 1;  // a plain store: every writer stores 1
}

// Eight waves: waves 0-3 own the first ceil(NB/2) blocks of SNP groups
// 4 blockIdx.x .. +3, waves 4-7 the rest of the same groups (and build the
// P buffers), so that two waves share each SIMD.
template <int NB, bool RAW>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_sketch_dense(
    const uint8_t* __restrict__ slabs, uint64_t pitch, uint32_t row_dw, uint32_t n,
    const uint64_t* __restrict__ bed_idx, uint64_t base, uint32_t m, uint32_t m_pad,
    const uint16_t* __restrict__ pcode, const uint8_t* __restrict__ miss_row,
    const int32_t* __restrict__ tsum, uint32_t* __restrict__ stats_out,
    double* __restrict__ norm, double* __restrict__ asum, void* __restrict__ cols,
    int32_t* __restrict__ flag, uint32_t d, int32_t* __restrict__ sg_out,
    int32_t* __restrict__ sm_out) {
    extern __shared__ uint32_t lds[];  // three phase buffers of P
    constexpr int NB0 = (NB + 1) / 2;
    if (threadIdx.x < 256)
        sketch_dense_half<NB, NB0, false, false, RAW>(
            lds, slabs, pitch, row_dw, n, bed_idx, base, m, m_pad, pcode, miss_row, tsum,
            stats_out, norm, asum, cols, flag, d, sg_out, sm_out, 0, 0);
    else
        sketch_dense_half<NB, NB - NB0, true, true, RAW>(
            lds, slabs, pitch, row_dw, n, bed_idx, base, m, m_pad, pcode, miss_row, tsum,
            stats_out, norm, asum, cols, flag, d, sg_out, sm_out, NB0, 1);
}

hipError_t launch_sketch_fused(const SketchPlan* sp, const ldsc_panel* panel,
                               const uint64_t* d_bed_idx, uint64_t base, uint32_t m, uint32_t m_pad,
                               uint32_t* stats_out, double* norm, void* cols, hipStream_t st,
                               int32_t* sg_out, int32_t* sm_out) {
    const bool raw = sg_out != nullptr;
    const uint32_t nb = (uint32_t)sp->nblocks, bpw = sketch_fused_bpw(nb);
    const uint32_t nranges = sketch_fused_ranges(nb), ngroups = m_pad / 32;
    const uint32_t per_xcd = (ngroups * nranges + 7) / 8;
    double* asum = raw ? nullptr : norm + (size_t)sketch_slots(sp) * m_pad;
    int32_t* flag = raw ? nullptr : sketch_i8_flag(sp, norm, m_pad);
    hipError_t e;
    if (!raw) {
        hipLaunchKernelGGL(k_sketch_clear_flag, dim3(1), dim3(64), 0, st, flag);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (sp->d_pcode) {  // natural order, at most kSdMaxBlocks blocks
        static const void* const dfn[2][kSdMaxBlocks] = {
            {(const void*)k_sketch_dense<1, false>, (const void*)k_sketch_dense<2, false>,
             (const void*)k_sketch_dense<3, false>, (const void*)k_sketch_dense<4, false>,
             (const void*)k_sketch_dense<5, false>, (const void*)k_sketch_dense<6, false>,
             (const void*)k_sketch_dense<7, false>},
            {(const void*)k_sketch_dense<1, true>, (const void*)k_sketch_dense<2, true>,
             (const void*)k_sketch_dense<3, true>, (const void*)k_sketch_dense<4, true>,
             (const void*)k_sketch_dense<5, true>, (const void*)k_sketch_dense<6, true>,
             (const void*)k_sketch_dense<7, true>}};
        const void* kfn = dfn[raw ? 1 : 0][nb - 1];
        const size_t lds = (size_t)3 * kSdPhasePairs * 2 * nb * 64 * 4 * 4;
        const uint8_t* slabs = panel->d_slabs;
        uint64_t pitch = panel->stride, a_base = base;
        uint32_t row_dw = (uint32_t)(panel->stride / 4), n = (uint32_t)panel->n;
        uint32_t a_m = m, a_mp = m_pad, a_d = (uint32_t)sp->d;
        const uint16_t* pc = sp->d_pcode;
        const uint8_t* miss = sp->d_miss;
        const int32_t* ts = sp->d_tsum;
        void* kargs[] = {(void*)&slabs, (void*)&pitch, (void*)&row_dw, (void*)&n,
                         (void*)&d_bed_idx, (void*)&a_base, (void*)&a_m, (void*)&a_mp,
                         (void*)&pc, (void*)&miss, (void*)&ts, (void*)&stats_out, (void*)&norm, (void*)&asum,
                         (void*)&cols, (void*)&flag, (void*)&a_d, (void*)&sg_out, (void*)&sm_out};
        e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        return hipLaunchKernel(kfn, dim3((ngroups + kSdWaves - 1) / kSdWaves),
                               dim3(2 * 64 * kSdWaves), kargs, lds, st);
    }
    const size_t lds = (size_t)(8 * kSfStride + kSfWaves * kSfStage) * 4;
    static const void* const fns[2][3] = {
        {(const void*)k_sketch_fused<1, false>, (const void*)k_sketch_fused<2, false>,
         (const void*)k_sketch_fused<4, false>},
        {(const void*)k_sketch_fused<1, true>, (const void*)k_sketch_fused<2, true>,
         (const void*)k_sketch_fused<4, true>}};
    const void* kfn = fns[raw ? 1 : 0][bpw == 1 ? 0 : bpw == 2 ? 1 : 2];
    e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint8_t* slabs = panel->d_slabs;
    uint64_t pitch = panel->stride, a_base = base;
    uint32_t src_dw = (uint32_t)((panel->bps + 3) / 4), n = (uint32_t)panel->n;
    const uint16_t* local = sp->d_local;
    uint32_t nch = (uint32_t)sp->nchunks, a_m = m, a_mp = m_pad;
    const v4i* pt = (const v4i*)sp->d_ptiles;
    const uint32_t* bo = sp->d_bo;
    const int32_t* bst = sp->d_bstart;
    const int32_t* ts = sp->d_tsum;
    uint32_t a_nb = nb, a_nr = nranges, a_ng = ngroups, a_px = per_xcd, a_d = (uint32_t)sp->d;
    void* kargs[] = {(void*)&slabs, (void*)&pitch, (void*)&src_dw, (void*)&n, (void*)&d_bed_idx,
                     (void*)&a_base, (void*)&a_m, (void*)&a_mp, (void*)&local, (void*)&nch,
                     (void*)&pt, (void*)&bo, (void*)&bst, (void*)&a_nb, (void*)&a_nr, (void*)&a_ng,
                     (void*)&a_px, (void*)&ts, (void*)&stats_out, (void*)&norm, (void*)&asum,
                     (void*)&cols, (void*)&flag, (void*)&a_d, (void*)&sg_out, (void*)&sm_out};
    return hipLaunchKernel(kfn, dim3(8 * per_xcd), dim3(64 * kSfWaves), kargs, lds, st);
}

hipError_t launch_sketch_fused_scale(const SketchPlan* sp, const float2* musc, uint32_t m_pad,
                                     double n, double* norm, hipStream_t st) {
    // k_sketch_dense writes one share slot, k_sketch_fused one per wave and
    // range
    const uint32_t slots = sp->d_pcode ? 2u : kSfWaves * sketch_fused_ranges(sp->nblocks);
    double* asum = norm + (size_t)sketch_slots(sp) * m_pad;
    double* consts = const_cast<double*>(sketch_i8_consts(sp, norm, m_pad));
    hipLaunchKernelGGL(k_sketch_scale<kSketchI8>, dim3((m_pad + 255) / 256), dim3(256), 0, st,
                       norm, slots, m_pad, n, nullptr, musc, asum, consts);
    return hipGetLastError();
}

// Three bf16 parts of an f32: y = hi + mid + lo + e with each part the
// round-to-nearest-even bf16 of what the previous ones leave (the
// subtractions are exact in f32), |mid| <= 2^-8 |y|, |lo| <= 2^-16 |y| and
// |e| <= 2^-25 |y|.
struct Bf16x3 {
    bf16x8 hi, mid, lo;
};

__device__ __forceinline__ Bf16x3 split3(const float4& p, const float4& q) {
    const float v[8] = {p.x, p.y, p.z, p.w, q.x, q.y, q.z, q.w};
    Bf16x3 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const __bf16 h = (__bf16)v[i];
        const float r1 = v[i] - (float)h;
        const __bf16 m = (__bf16)r1;
        o.hi[i] = h;
        o.mid[i] = m;
        o.lo[i] = (__bf16)(r1 - (float)m);
    }
    return o;
}

// One 16-bucket step of the 64x64 tile: the six part products of order
// <= 2, smallest first, the four accumulators taking turns so that
// consecutive MFMAs are independent.
__device__ __forceinline__ void sk_mma(const Bf16x3 (&op)[4], v16f (&acc)[2][2]) {
#pragma unroll
    for (int pp = 0; pp < 6; ++pp)
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
                const Bf16x3& A = op[bi];
                const Bf16x3& Bo = op[2 + bj];
                const bf16x8& pa = pp == 0 ? A.lo : pp == 1 || pp == 3 ? A.mid : A.hi;
                const bf16x8& pb = pp == 2 ? Bo.lo : pp == 1 || pp == 4 ? Bo.mid : Bo.hi;
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pa, pb, acc[bi][bj], 0, 0, 0);
            }
}

// Asks the scheduler to issue the split's VALU work between the 24 MFMAs
// of a step rather than before them.
__device__ __forceinline__ void sk_interleave() {
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
        __builtin_amdgcn_sched_group_barrier(0x002, 10, 0); // then VALU
    }
    __builtin_amdgcn_sched_barrier(0);
}

// d-length contraction of the sketched columns over one 64x64 band tile on
// the bf16 MFMA.  Each f32 value is split into three bf16 parts in
// registers and a product is formed from the six part products of order
// <= 2 (hi*hi, hi*mid, mid*hi, hi*lo, mid*mid, lo*hi), smallest first; the
// dropped ones are below 2^-23 of it, so each product is within about
// 2^-22 relative, next to the f32 rounding of the FP32 MFMA it replaces,
// at six 32x32x16 bf16 steps per 16 buckets instead of eight 32x32x2 f32
// steps.  Value r of a lane is slot r % 8 of step r / 8 for both operands,
// so each step adds the products of the same 16 buckets.
__global__ __launch_bounds__(64) void k_band_sk(BandArgs a) {
    __shared__ EpiShared sh;
    __shared__ double kr[64], kc[64], scr[64], scc[64];
    const int tile = tile_of_block(a);
    if (tile < 0) return;
    const int2 tl = a.tiles[tile];
    const uint32_t I = (uint32_t)tl.x, J = (uint32_t)tl.y;
    const int lane = threadIdx.x;
    const size_t gstride = (size_t)a.nkb * 64 * 4;  // float4 per group
    const float4* pa0 = reinterpret_cast<const float4*>(a.sk) + (size_t)(2 * I) * gstride + lane * 4;
    const float4* pa1 = pa0 + gstride;
    const float4* pb0 = reinterpret_cast<const float4*>(a.sk) + (size_t)(2 * J) * gstride + lane * 4;
    const float4* pb1 = pb0 + gstride;

    v16f acc[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[bi][bj][r] = 0.0f;

    // x[operand][q]: values 4q .. 4q+3 of the lane; operands A0, A1, B0, B1.
    // Software pipeline over the two 16-bucket steps of each block: the
    // split of the next step runs on the VALU while the MFMAs of the current
    // one run (e0 holds step 0 of a block, e1 step 1).
    float4 x[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        x[0][q] = pa0[q];
        x[1][q] = pa1[q];
        x[2][q] = pb0[q];
        x[3][q] = pb1[q];
    }
    Bf16x3 e0[4], e1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) e0[t] = split3(x[t][0], x[t][1]);
    for (uint32_t B = 0; B < a.nkb; ++B) {
        const size_t o = (size_t)(B + 1 < a.nkb ? B + 1 : B) * 256;
        float4 y[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            y[0][q] = pa0[o + q];
            y[1][q] = pa1[o + q];
            y[2][q] = pb0[o + q];
            y[3][q] = pb1[o + q];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch a block ahead
#pragma unroll
        for (int t = 0; t < 4; ++t) e1[t] = split3(x[t][2], x[t][3]);
        sk_mma(e0, acc);
        sk_interleave();
#pragma unroll
        for (int t = 0; t < 4; ++t) e0[t] = split3(y[t][0], y[t][1]);
        sk_mma(e1, acc);
        sk_interleave();
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) x[t][q] = y[t][q];
    }

    stage_info(a, sh, I, J);
    kr[lane] = a.kappa ? a.kappa[(size_t)I * kTile + lane] : 0.0;
    kc[lane] = a.kappa ? a.kappa[(size_t)J * kTile + lane] : 0.0;
    scr[lane] = (double)a.sk_scale[(size_t)I * kTile + lane];
    scc[lane] = (double)a.sk_scale[(size_t)J * kTile + lane];
    __syncthreads();
    // quadratic_sketch_correction[_maf_aware] (compute.rs:947-998)
    double r2u[2][2][16];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = row_of(bi, r, lane), col = 32 * bj + (lane & 31);
                const double v = (double)acc[bi][bj][r] * scr[row] * scc[col];
                const double t = v * v * a.n_inv_sq;
                const double dm3 = a.d_f - 3.0 - 2.0 * (kr[row] + kc[col]);
                const double disc = dm3 * dm3 + 8.0 * (a.d_f * t - 1.0);
                const double rs = disc >= 0.0 ? (-dm3 + sqrt(disc)) * 0.25 : 0.0;
                const double val = rs * a.r2u_a + a.r2u_b;
                r2u[bi][bj][r] = pair_included(a, sh, I, J, row, col) ? val : 0.0;
            }
    band_contract(a, sh, tile, I, J, r2u);
}

hipError_t launch_band_sk(const BandArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_band_sk, dim3(8 * a.per_xcd), dim3(64), 0, st, a);
    return hipGetLastError();
}

// d-length contraction of the integer-operand form over one 64x64 band tile
// on the int8 MFMA: with a = w + delta e per bucket, the sketched dot
// product of SNPs r and c is (quadratic_sketch_correction's val before the
// 1/N^2 of compute.rs:947-998)
//   v = scale_r scale_c [ W_r.W_c + d_c W_r.E_c + d_r E_r.W_c + d_r d_c E_r.E_c
//                         - d_c A_r - d_r A_c + d_r d_c TT ],
// four exact int32 products over the d_pad buckets (32 per K-step), finished
// in FP64.  The next K-steps' operands load while the current ones' MFMAs
// run.
constexpr int kSk8Ahead = 3;

// Square root of x > 0 in FP64 for arguments far from the subnormal and
// overflow ranges (here the discriminant of the bias inversion, of the
// order of d^2): v_rsq_f64 refined by the coupled Newton-Raphson iteration
// for sqrt(x) and 1/(2 sqrt(x)) (Markstein's method: one step on both, then
// two corrections of g from the residual x - g^2), so the result is within
// an ulp of the rounded root.  The library sqrt adds range scaling and
// special-case tests this argument does not need.
__device__ __forceinline__ double sqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __fma_rn(-h, g, 0.5);
    g = __fma_rn(g, r, g);
    h = __fma_rn(h, r, h);
    double dd = __fma_rn(-g, g, x);
    g = __fma_rn(dd, h, g);
    dd = __fma_rn(-g, g, x);
    return __fma_rn(dd, h, g);
}

// Two waves per tile: wave w owns rows 32w .. 32w+31 (SNP group 2I + w)
// against both column groups, 4 products x 2 blocks of int32 accumulators
// (128 registers), so two waves fit per SIMD and one tile's FP64 epilogue
// overlaps another's MFMAs (measured faster than one wave per tile with all
// four blocks at d = 200, 1000 and 5000).  Each wave loads W and E of its
// rows and of one column group, 2J + w, and passes that group's operands to
// the other wave through LDS (two buffers, one barrier per K-step), so the
// tile's column operands cross the cache hierarchy once.  A wave's
// accumulator block 0 is its own column group (static register indices),
// block 1 the other wave's; the epilogue puts them back in column order.
//
// Epilogue per pair, the inversion of quadratic_sketch_correction
// (compute.rs:947-998) rearranged to fewer FP64 operations: with
// u = D scale_r scale_c sqrt(8 d / N^2), the discriminant is
// dm3^2 + 8 (d t - 1) = u^2 + (dm3^2 - 8), and
// r2u = rs r2u_a + r2u_b = sqrt(disc) (r2u_a / 4) + (r2u_b - dm3 r2u_a / 4)
// (r2u_b when disc < 0).  Without the MAF-aware kurtosis, dm3 = d - 3 and
// both constants are per launch.
template <bool MAF>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_band_sk8_2w(BandArgs a) {
    __shared__ EpiShared2 sh;
    __shared__ double kr[64], kc[64], scr[64], scc[64], dlr[64], dlc[64], ar[64], ac[64];
    const int tile = tile_of_block(a);
    if (tile < 0) return;
    const int2 tl = a.tiles[tile];
    const uint32_t I = (uint32_t)tl.x, J = (uint32_t)tl.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t gs = (size_t)a.nkb * 64;
    // operand streams: W and E of the wave's row group and of its column group
    const u32x4* src[4] = {a.skw + (2 * I + w) * gs + lane, a.ske + (2 * I + w) * gs + lane,
                           a.skw + (2 * J + w) * gs + lane, a.ske + (2 * J + w) * gs + lane};
    // the column operands exchanged: xch[buffer][wave][W or E][lane], in the
    // epilogue's LDS (unused until the K loop is done)
    auto xch = reinterpret_cast<u32x4 (*)[2][2][64]>(&sh.u);
    v16i acc[4][2];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[p][bj][r] = 0;
    u32x4 ring[kSk8Ahead][4];
    const uint32_t nk = a.nkb;
#pragma unroll
    for (int u = 0; u < kSk8Ahead; ++u) {
        const size_t o = (size_t)(u < (int)nk ? u : nk - 1) * 64;
#pragma unroll
        for (int q = 0; q < 4; ++q) ring[u][q] = src[q][o];
    }
    // every branch below is uniform over the workgroup (nk, kb, u), so both
    // waves reach each barrier
    for (uint32_t kb = 0; kb < nk; kb += kSk8Ahead) {
#pragma unroll
        for (int u = 0; u < kSk8Ahead; ++u) {
            if (kb + u < nk) {
                const uint32_t step = kb + u, buf = step & 1u;
                v4i op[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) op[q] = __builtin_bit_cast(v4i, ring[u][q]);
                xch[buf][w][0][lane] = ring[u][2];
                xch[buf][w][1][lane] = ring[u][3];
                const uint32_t nxt = step + kSk8Ahead;
                const size_t o = (size_t)(nxt < nk ? nxt : nk - 1) * 64;
#pragma unroll
                for (int q = 0; q < 4; ++q) ring[u][q] = src[q][o];
                // buffer `buf` was last read at step - 2, before the barrier
                // of step - 1
                __syncthreads();
                const v4i ow = __builtin_bit_cast(v4i, xch[buf][w ^ 1][0][lane]);
                const v4i oe = __builtin_bit_cast(v4i, xch[buf][w ^ 1][1][lane]);
                acc[0][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[0], op[2], acc[0][0], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[0], op[3], acc[1][0], 0, 0, 0);
                acc[2][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[1], op[2], acc[2][0], 0, 0, 0);
                acc[3][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[1], op[3], acc[3][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[0], ow, acc[0][1], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[0], oe, acc[1][1], 0, 0, 0);
                acc[2][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[1], ow, acc[2][1], 0, 0, 0);
                acc[3][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(op[1], oe, acc[3][1], 0, 0, 0);
            }
        }
    }
    // per-SNP constants: wave 0 the tile's rows (scale times sqrt(8 d / N^2)),
    // wave 1 its columns
    if (w == 0) {
        const size_t ri = (size_t)I * kTile + lane;
        sh.lor[lane] = a.lo[ri];
        if (MAF) kr[lane] = a.kappa[ri];
        scr[lane] = a.sk8[ri] * sqrt(8.0 * a.d_f * a.n_inv_sq);
        dlr[lane] = a.sk8[a.m_pad + ri];
        ar[lane] = a.sk8[2 * (size_t)a.m_pad + ri];
    } else {
        const size_t ci = (size_t)J * kTile + lane;
        if (MAF) kc[lane] = a.kappa[ci];
        scc[lane] = a.sk8[ci];
        dlc[lane] = a.sk8[a.m_pad + ci];
        ac[lane] = a.sk8[2 * (size_t)a.m_pad + ci];
    }
    __syncthreads();
    const double qa = 0.25 * a.r2u_a;
    const double dm3_0 = a.d_f - 3.0;
    const double c0_0 = __fma_rn(dm3_0, dm3_0, -8.0);
    const double bb_0 = __fma_rn(-dm3_0, qa, a.r2u_b);
    double val2[2][16];  // in accumulator order: block bj is column block bj ^ w
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 32 * w + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = 32 * (bj ^ w) + (lane & 31);
            const double dr = dlr[row], dc = dlc[col];
            const double D = (double)acc[0][bj][r] + dc * (double)acc[1][bj][r] +
                             dr * (double)acc[2][bj][r] +
                             dr * dc * ((double)acc[3][bj][r] + a.sk_tt) - dc * ar[row] - dr * ac[col];
            const double u = D * scr[row] * scc[col];
            double c0 = c0_0, bb = bb_0;
            if (MAF) {
                const double dm3 = a.d_f - 3.0 - 2.0 * (kr[row] + kc[col]);
                c0 = __fma_rn(dm3, dm3, -8.0);
                bb = __fma_rn(-dm3, qa, a.r2u_b);
            }
            const double disc = __fma_rn(u, u, c0);
            const double sq = disc > 0.0 ? sqrt_pos(disc) : 0.0;
            const double val = disc >= 0.0 ? __fma_rn(sq, qa, bb) : a.r2u_b;
            const uint32_t jj = I * kTile + row, ii = J * kTile + col;
            val2[bj][r] = (ii < jj && jj < a.m && ii >= sh.lor[row]) ? val : 0.0;
            // with the kurtosis, pair by pair: hoisted reads would spill
            if (MAF) __builtin_amdgcn_sched_barrier(0);
        }
    double r2u[2][16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        r2u[0][r] = w ? val2[1][r] : val2[0][r];
        r2u[1][r] = w ? val2[0][r] : val2[1][r];
    }
    band_contract2(a, sh, tile, I, J, w, r2u);
}

hipError_t launch_band_sk8(const BandArgs& a, hipStream_t st) {
    if (a.kappa)
        hipLaunchKernelGGL(k_band_sk8_2w<true>, dim3(8 * a.per_xcd), dim3(128), 0, st, a);
    else
        hipLaunchKernelGGL(k_band_sk8_2w<false>, dim3(8 * a.per_xcd), dim3(128), 0, st, a);
    return hipGetLastError();
}

}  // namespace ldsc

using namespace ldsc;

extern "C" int ldsc_l2_sketch_buckets(const ldsc_panel* panel, const uint64_t* bed_idx, uint64_t m,
                                      uint64_t d, int32_t* sg, int32_t* sm, char* err,
                                      uint64_t err_len) {
    if (!panel) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "panel is NULL");
    if (d < 1 || d > 1000000)
        return fail(err, err_len, LDSC_ERR_SKETCH_RANGE, "sketch dimension %llu out of range",
                    (unsigned long long)d);
    if (m == 0) return LDSC_OK;
    if (!bed_idx || !sg || !sm) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL array");
    for (uint64_t j = 0; j < m; ++j)
        if (bed_idx[j] >= panel->m_bim)
            return fail(err, err_len, LDSC_ERR_INVALID_ARG, "bed_idx %llu out of range",
                        (unsigned long long)bed_idx[j]);
    if (m >= (1ull << 31)) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "too many SNPs");
    Workspace* ws = acquire_workspace(panel->device);
    if (!ws) return fail(err, err_len, LDSC_ERR_DEVICE, "could not create a device workspace");
    int rc = LDSC_OK;
    do {
        SketchPlan* sp = nullptr;
        hipError_t e;
#define SCHK(x) if ((e = (x)) != hipSuccess) { rc = fail(err, err_len, e == hipErrorOutOfMemory ? LDSC_ERR_OOM : LDSC_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e)); break; }
        SCHK(hipSetDevice(panel->device));
        rc = sketch_plan(panel, d, &sp, err, err_len);
        if (rc) break;
        const uint64_t m_pad = round_up(m, kTile);
        hipStream_t st = ws->stream;
        SCHK(ws->bed_idx.ensure(m * 8));
        SCHK(ws->info.ensure(m_pad * 48));
        SCHK(ws->sk_cols.ensure(m * d * 8));
        SCHK(hipMemcpyAsync(ws->bed_idx.p, bed_idx, m * 8, hipMemcpyHostToDevice, st));
        int32_t* d_sg = ws->sk_cols.as<int32_t>();
        int32_t* d_sm = d_sg + m * d;
        if (sp->d_local && sp->d_pad <= kSketchI8MaxDpad) {
            // (every block count here, so that the kernel's forms with several
            // blocks per wave stay checked bit for bit)
            // chunk-local order: the one-pass kernel's exact sums
            SCHK(ws->stats.ensure(m_pad * 16));
            SCHK(launch_sketch_fused(sp, panel, ws->bed_idx.as<uint64_t>(), 0, (uint32_t)m,
                                     (uint32_t)m_pad, ws->stats.as<uint32_t>(), nullptr, nullptr,
                                     st, d_sg, d_sm));
            SCHK(hipMemcpyAsync(sg, d_sg, m * d * 4, hipMemcpyDeviceToHost, st));
            SCHK(hipMemcpyAsync(sm, d_sm, m * d * 4, hipMemcpyDeviceToHost, st));
            SCHK(hipStreamSynchronize(st));
            break;
        }
        uint32_t parts = 0;
        rc = sp->d_local ? gather_tile_chunks(panel, ws->bed_idx.as<uint64_t>(), 0, m,
                                              sp->d_local, (uint32_t)sp->chunk, m_pad, ws,
                                              &parts, err, err_len)
                         : gather_and_tile(panel, ws->bed_idx.as<uint64_t>(), 0, m, sp->d_fam,
                                           m_pad, true, /*tile=*/true, ws, &parts, err, err_len);
        if (rc) break;
        // musc is unused by the integer form; pass any valid per-SNP array
        SCHK(hipMemsetAsync(ws->info.p, 0, m_pad * 48, st));
        const uint32_t nkb = (uint32_t)(round_up(panel->n, kKBlock) / kKBlock);
        SCHK(launch_sketch(sp, kSketchRaw, ws->tiled.as<uint32_t>(), nkb,
                           reinterpret_cast<const float2*>(ws->info.as<char>()), (uint32_t)m,
                           (uint32_t)m_pad, (double)panel->n, nullptr, nullptr, d_sg, d_sm, st));
        SCHK(hipMemcpyAsync(sg, d_sg, m * d * 4, hipMemcpyDeviceToHost, st));
        SCHK(hipMemcpyAsync(sm, d_sm, m * d * 4, hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
#undef SCHK
    } while (0);
    release_workspace(ws);
    return rc;
}
