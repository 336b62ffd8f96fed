// band_common.h -- device code shared by the exact and the sketch band
// kernels: the tile list, the XCD-aware tile mapping, the masked r2u
// epilogue and the deterministic per-tile annotation contraction.
#pragma once

#include <algorithm>

#include "device.h"

namespace ldsc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// v_mfma_scale_f32_32x32x64_f8f6f4 with e2m1 (FP4) operands and scale bytes
// 127 (E8M0 2^0) on both: a plain e2m1 product with f32 accumulation
#define LDSC_MFMA_FP4(A, B, C) \
    __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(A, B, C, 4, 4, 0, 0x7f7f7f7f, 0, 0x7f7f7f7f)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------------
// Kernel 3: banded contraction (BandArgs: device.h)
// ------------------------------------------------------------------------

// Tile -> XCD mapping: workgroups are dealt to the 8 XCDs round-robin, so
// workgroup b runs on XCD b%8.  Give each XCD one contiguous stretch of the
// (I, J)-sorted tile list: neighbouring tiles share SNP groups, which then
// hit in that XCD's L2.
__device__ inline int tile_of_block(const BandArgs& a) {
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const uint32_t t = xcd * a.per_xcd + slot;
    return t < a.ntiles ? (int)(a.tile0 + t) : -1;
}

// Several calls' bands in one launch (the request server): the tiles of
// call r are global tiles starts[r] .. starts[r+1]-1, each with the call's
// own arguments args[r] (device memory).  The same XCD-aware stretches as
// tile_of_block are laid over the concatenated tile list.
struct BandBatch {
    const BandArgs* args;
    const uint32_t* starts;  // nreq + 1 cumulative tile counts
    uint32_t nreq, ntiles, per_xcd;
};

__device__ inline bool batch_tile(const BandBatch& b, uint32_t& r, int& tile) {
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const uint32_t t = xcd * b.per_xcd + slot;
    if (t >= b.ntiles) return false;
    uint32_t lo = 0, hi = b.nreq;  // starts[lo] <= t < starts[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (b.starts[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    r = lo;
    tile = (int)(t - b.starts[lo]);
    return true;
}

// Expand one dword of 16 PLINK codes (device.h bit order) into int8 dosage
// G (00->2, 01->0, 10->1, 11->0) and presence M (01->0, else 1).  Each
// byte of D holds one code 0..3; v_perm_b32 uses it as a byte selector into
// the 4-byte table.  Both perm sources carry the table, so selectors 0..3
// and 4..7 read the same entry.
__device__ __forceinline__ void expand_codes(uint32_t w, v4i& G, v4i& M) {
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const uint32_t D = (w >> (2 * f)) & 0x03030303u;
        G[f] = (int)__builtin_amdgcn_perm(0x00010002u, 0x00010002u, D);
        M[f] = (int)__builtin_amdgcn_perm(0x01010001u, 0x01010001u, D);
    }
}

// e2m1 operands of one K-step from the lane's two value-code dwords w0, w1
// (32 individuals): X = v/2 (nibble = v in its low bits: 0, 0.5, 1, 1.5) and
// Q = q/2 (q = 1 iff v = 3: missing or padding).  Fields 2n and 2n+1 of a
// dword go to nibble n of two operand dwords; the same placement on both
// operands keeps every individual paired with itself.  The MFMA reads 4 of
// the 8 operand dwords.
__device__ __forceinline__ void vcode_operands(uint32_t w0, uint32_t w1, v8i& X, v8i& Q) {
    // shifts of the 64-bit pair (w1:w0): the bits w1 pushes into the top of
    // w0 land outside every mask below
    const uint64_t w = ((uint64_t)w1 << 32) | w0;
    const uint64_t s1 = w >> 1, s2 = w >> 2, s3 = w >> 3;
    const uint32_t a2 = (uint32_t)s2, b2 = (uint32_t)(s2 >> 32);
    const v4i x = {(int)(w0 & 0x33333333u), (int)(a2 & 0x33333333u),
                   (int)(w1 & 0x33333333u), (int)(b2 & 0x33333333u)};
    const v4i q = {(int)(w0 & (uint32_t)s1 & 0x11111111u), (int)(a2 & (uint32_t)s3 & 0x11111111u),
                   (int)(w1 & (uint32_t)(s1 >> 32) & 0x11111111u),
                   (int)(b2 & (uint32_t)(s3 >> 32) & 0x11111111u)};
    X = __builtin_shufflevector(x, x, 0, 1, 2, 3, -1, -1, -1, -1);
    Q = __builtin_shufflevector(q, q, 0, 1, 2, 3, -1, -1, -1, -1);
}

// Cells of individuals >= nv (the row's real count; 0 for a padding row)
// of dword q read as code 01 (missing).
__device__ __forceinline__ uint32_t ts_mask(uint32_t raw, uint64_t q, uint32_t nv) {
    const int64_t left = (int64_t)nv - (int64_t)(q * 16);
    const uint32_t keep = left >= 16 ? 0xFFFFFFFFu
                        : left <= 0  ? 0u
                                     : (1u << (2 * (uint32_t)left)) - 1u;
    return (raw & keep) | (0x55555555u & ~keep);
}

// o[k] byte r = byte k of d_r (4x4 byte transpose).
__device__ __forceinline__ void transpose_bytes4(uint32_t d0, uint32_t d1, uint32_t d2,
                                                 uint32_t d3, uint32_t (&o)[4]) {
    const uint32_t lo01 = __builtin_amdgcn_perm(d1, d0, 0x05010400u);  // d0.b0 d1.b0 d0.b1 d1.b1
    const uint32_t hi01 = __builtin_amdgcn_perm(d1, d0, 0x07030602u);  // d0.b2 d1.b2 d0.b3 d1.b3
    const uint32_t lo23 = __builtin_amdgcn_perm(d3, d2, 0x05010400u);
    const uint32_t hi23 = __builtin_amdgcn_perm(d3, d2, 0x07030602u);
    o[0] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    o[1] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    o[2] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    o[3] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// The reference's f32 mean and inverse std of a SNP from its exact integer
// stats (normalize.rs:115-154): the same operation order, no contraction
// into FMA.  Returns the f64 mean.
__device__ __forceinline__ double ref_mean_inv_f32(uint32_t S, uint32_t C, uint32_t SS, uint32_t n,
                                                   float& a32, float& i32) {
    const double avg = C > 0 ? (double)S / (double)C : 0.0;
    const double css = __dsub_rn((double)SS, __dmul_rn(__dmul_rn((double)C, avg), avg));
    const double sd = sqrt(css / (double)n);
    a32 = (float)avg;
    i32 = sd > 0.0 ? (float)(1.0 / sd) : 0.0f;
    return avg;
}

struct EpiShared {
    union {
        struct {              // one annotation: one 32-row half of R at a time
            double R[32][65];
            double ar[64];    // A_eff of the tile's row SNPs
            double ac[64];    // A_eff of the tile's column SNPs
        } one;
        double Rf[64][65];    // several annotations: the whole tile's masked r2u
    } u;
    int2 csr[64], csc[64];
    double rsr[64], rsc[64];
    uint32_t lor[64];
};

__device__ __forceinline__ int row_of(int bi, int r, int lane) {
    return 32 * bi + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// SNP stride of A_eff in HBM (m_pad rows of aeff_stride(K) doubles,
// annotations contiguous): 1 for a single annotation, else K rounded up to
// the 16 annotations one FP64 MFMA pass covers (padding columns are zero).
__host__ __device__ inline uint32_t aeff_stride(uint32_t K) {
    return K == 1 ? 1u : (K + 15u) & ~15u;
}

// One annotation: r2u (masked, 64 per lane in MFMA output order) ->
//   rowp[tile][0][a] = sum_b R[a][b] * A_eff[col b]
//   colp[tile][0][b] = sum_a R[a][b] * A_eff[row a]
// on the vector FP64 units, one 32-row half of R through LDS at a time.
__device__ __forceinline__ void band_contract_one(const BandArgs& a, EpiShared& sh, int tile,
                                                  uint32_t I, uint32_t J,
                                                  const double (&r2u)[2][2][16]) {
    const int lane = threadIdx.x;
    __syncthreads();
    sh.u.one.ar[lane] = a.aeff[(size_t)I * kTile + lane];
    sh.u.one.ac[lane] = a.aeff[(size_t)J * kTile + lane];
    double cacc = 0.0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        __syncthreads();
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                sh.u.one.R[row_of(0, r, lane)][32 * bj + (lane & 31)] = r2u[bi][bj][r];
        __syncthreads();
        // row role: lane covers row lane%32 over columns of its half
        double racc = 0.0;
        const int rr = lane & 31, b0 = 32 * (lane >> 5);
        for (int b = 0; b < 32; ++b) racc += sh.u.one.R[rr][b0 + b] * sh.u.one.ac[b0 + b];
        racc += __shfl_xor(racc, 32);
        if (lane < 32) a.rowp[(size_t)tile * kTile + 32 * bi + lane] = racc;
        // column role: lane = column, rows of this half
        for (int r = 0; r < 32; ++r) cacc += sh.u.one.R[r][lane] * sh.u.one.ar[32 * bi + r];
    }
    a.colp[(size_t)tile * kTile + lane] = cacc;
}

typedef double v4d __attribute__((ext_vector_type(4)));

// Several annotations: R (the tile's masked r2u, row-major in LDS) against
// the A_eff rows of the tile's SNPs on the FP64 MFMA, 16 annotations per
// pass:
//   ROWS: rowp[tile][k][a] = sum_b R[a][b] * A_eff[col b][k]
//   COLS: colp[tile][k][b] = sum_a R[a][b] * A_eff[row a][k]
// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and holds D[4i + l/16][l%16] in its register i = 0..3.  Each 16x16 output block sums
// its 64 terms in a fixed order (16 steps of 4), so results are
// reproducible and equal whichever wave computes them.
template <bool ROWS, bool COLS>
__device__ __forceinline__ void band_contract_mfma(const BandArgs& a, const double (*Rf)[65],
                                                   int tile, uint32_t I, uint32_t J) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const uint32_t Kp = aeff_stride(a.K);
    const double* ac = a.aeff + (size_t)(J * kTile + lk) * Kp + li;
    const double* ar = a.aeff + (size_t)(I * kTile + lk) * Kp + li;
    for (uint32_t k0 = 0; k0 < a.K; k0 += 16) {
        double bc[16], br[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            if (ROWS) bc[s] = ac[(size_t)4 * s * Kp + k0];
            if (COLS) br[s] = ar[(size_t)4 * s * Kp + k0];
        }
        v4d dr[4], dc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            dr[q] = v4d{0.0, 0.0, 0.0, 0.0};
            dc[q] = v4d{0.0, 0.0, 0.0, 0.0};
        }
        // A operands of step s+1 are read from LDS while step s's MFMAs
        // run; the barrier keeps the scheduler from hoisting further reads
        // (which would only add live registers)
        double xr[4], xc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xr[q] = ROWS ? Rf[16 * q + li][lk] : 0.0;
            xc[q] = COLS ? Rf[lk][16 * q + li] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            double yr[4], yc[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                yr[q] = xr[q];
                yc[q] = xc[q];
                if (s + 1 < 16) {
                    if (ROWS) xr[q] = Rf[16 * q + li][4 * (s + 1) + lk];
                    if (COLS) xc[q] = Rf[4 * (s + 1) + lk][16 * q + li];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (ROWS) dr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(yr[q], bc[s], dr[q], 0, 0, 0);
                if (COLS) dc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(yc[q], br[s], dc[q], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        const uint32_t k = k0 + li;
        if (k < a.K) {
            const size_t o = ((size_t)tile * a.K + k) * kTile + lk;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (ROWS) a.rowp[o + 16 * q + 4 * i] = dr[q][i];
                    if (COLS) a.colp[o + 16 * q + 4 * i] = dc[q][i];
                }
        }
    }
}

// r2u (masked, 64 per lane in MFMA output order) -> the tile's partial
// sums for every annotation (single-wave kernels).
__device__ __forceinline__ void band_contract(const BandArgs& a, EpiShared& sh, int tile,
                                              uint32_t I, uint32_t J,
                                              const double (&r2u)[2][2][16]) {
    if (a.K == 1) {
        band_contract_one(a, sh, tile, I, J, r2u);
        return;
    }
    const int lane = threadIdx.x;
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                sh.u.Rf[row_of(bi, r, lane)][32 * bj + (lane & 31)] = r2u[bi][bj][r];
    __syncthreads();
    band_contract_mfma<true, true>(a, sh.u.Rf, tile, I, J);
}

__device__ __forceinline__ void stage_info(const BandArgs& a, EpiShared& sh, uint32_t I,
                                           uint32_t J) {
    const int lane = threadIdx.x;
    sh.csr[lane] = a.cs[(size_t)I * kTile + lane];
    sh.rsr[lane] = a.rs[(size_t)I * kTile + lane];
    sh.lor[lane] = a.lo[(size_t)I * kTile + lane];
    sh.csc[lane] = a.cs[(size_t)J * kTile + lane];
    sh.rsc[lane] = a.rs[(size_t)J * kTile + lane];
}

__device__ __forceinline__ bool pair_included(const BandArgs& a, const EpiShared& sh,
                                              uint32_t I, uint32_t J, int row, int col) {
    const uint32_t jj = I * kTile + row, ii = J * kTile + col;
    return ii < jj && jj < a.m && ii >= sh.lor[row];
}

__device__ __forceinline__ float pick_code(const float4& L, uint32_t code) {
    const float lo = (code & 1u) ? L.y : L.x;
    const float hi = (code & 1u) ? L.w : L.z;
    return (code & 2u) ? hi : lo;
}


// Epilogue storage of the two-wave band kernel.  One annotation: each wave
// stages one 16-row quarter of its 32 x 64 r2u block at a time.  Several:
// the whole tile's r2u for the FP64 MFMA contraction.  ~35 KiB per
// workgroup, so four workgroups (two waves per SIMD) fit in a CU's LDS.
struct EpiShared2 {
    union {
        struct {
            double R[2][16][65];
            double ar[64];    // A_eff of the tile's row SNPs
            double ac[64];    // A_eff of the tile's column SNPs
            double cx[64];    // wave 1's column partials
        } one;
        double Rf[64][65];
    } u;
    int2 csr[64], csc[64];
    double rsr[64], rsc[64];
    uint32_t lor[64];
};

// One annotation: r2u of wave w (rows 32w .. 32w+31 of the tile, all 64
// columns, in MFMA output order r2u[bj][r]) -> the tile's partial sums,
// fixed summation order.
__device__ __forceinline__ void band_contract2_one(const BandArgs& a, EpiShared2& sh, int tile,
                                                   uint32_t I, uint32_t J, int w,
                                                   const double (&r2u)[2][16]) {
    const int lane = threadIdx.x & 63;
    __syncthreads();
    if (w == 0)
        sh.u.one.ar[lane] = a.aeff[(size_t)I * kTile + lane];
    else
        sh.u.one.ac[lane] = a.aeff[(size_t)J * kTile + lane];
    double cacc = 0.0;
#pragma unroll
    for (int Q = 0; Q < 2; ++Q) {
        __syncthreads();
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int r = 8 * Q + i;
                const int local = 8 * ((r >> 2) - 2 * Q) + 4 * (lane >> 5) + (r & 3);
                sh.u.one.R[w][local][32 * bj + (lane & 31)] = r2u[bj][r];
            }
        __syncthreads();
        // row role: row lane%16 of the quarter, columns 16*(lane/16) .. +16
        double racc = 0.0;
        const int rr = lane & 15, b0 = 16 * (lane >> 4);
        for (int b = 0; b < 16; ++b) racc += sh.u.one.R[w][rr][b0 + b] * sh.u.one.ac[b0 + b];
        racc += __shfl_xor(racc, 16);
        racc += __shfl_xor(racc, 32);
        if (lane < 16) a.rowp[(size_t)tile * kTile + 32 * w + 16 * Q + lane] = racc;
        // column role: lane = column, the 16 rows of this quarter
        for (int r = 0; r < 16; ++r)
            cacc += sh.u.one.R[w][r][lane] * sh.u.one.ar[32 * w + 16 * Q + r];
    }
    __syncthreads();
    if (w == 1) sh.u.one.cx[lane] = cacc;
    __syncthreads();
    if (w == 0) a.colp[(size_t)tile * kTile + lane] = cacc + sh.u.one.cx[lane];
}

__device__ __forceinline__ void band_contract2(const BandArgs& a, EpiShared2& sh, int tile,
                                               uint32_t I, uint32_t J, int w,
                                               const double (&r2u)[2][16]) {
    if (a.K == 1) {
        band_contract2_one(a, sh, tile, I, J, w, r2u);
        return;
    }
    // several annotations: both waves stage their rows, then wave 0 forms
    // the row partials and wave 1 the column partials
    const int lane = threadIdx.x & 63;
    __syncthreads();
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 16; ++r)
            sh.u.Rf[32 * w + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)][32 * bj + (lane & 31)] =
                r2u[bj][r];
    __syncthreads();
    if (w == 0)
        band_contract_mfma<true, false>(a, sh.u.Rf, tile, I, J);
    else
        band_contract_mfma<false, true>(a, sh.u.Rf, tile, I, J);
}

// The band's tiles: row block I pairs with column blocks J0(I) .. I, where
// J0 is the block of the pair bound of the block's first SNP (the bounds do
// not decrease).  plan_band finds the counts and starts; fill_band_tiles
// writes the list (ntiles entries) in parallel pieces.
inline void plan_band(const std::vector<uint64_t>& lo, uint64_t m, BandPlan& p) {
    p.m_pad = (uint32_t)round_up(m, kTile);
    p.nb = p.m_pad / kTile;
    p.j0.resize(p.nb);
    p.tstart.resize(p.nb + 1);
    p.tstart[0] = 0;
    for (uint32_t I = 0; I < p.nb; ++I) {
        const int J0 = (int)(lo[(uint64_t)I * kTile] / kTile);
        p.j0[I] = J0;
        p.tstart[I + 1] = p.tstart[I] + ((int)I - J0 + 1);
    }
    p.ntiles = (uint32_t)p.tstart[p.nb];
}

inline void fill_band_tiles(const BandPlan& p, int2* tiles) {
    const int T = task_count((uint64_t)p.ntiles);
    parallel_tasks(T, [&](int t) {
        const uint32_t a = (uint32_t)((uint64_t)p.nb * t / T), b = (uint32_t)((uint64_t)p.nb * (t + 1) / T);
        for (uint32_t I = a; I < b; ++I)
            for (int J = p.j0[I]; J <= (int)I; ++J)
                tiles[(size_t)p.tstart[I] + (J - p.j0[I])] = make_int2((int)I, J);
    });
}

// CountSketch (sketch.hip).  Forms of the bucket sums:
//  * kSketchRaw: the integer sums SG, SM (d x m column-major int32 each);
//  * kSketchF32: f32 bucket values in the lane order of k_band_sk, their
//    norm shares and per-SNP renorm scales;
//  * kSketchI8: the int8 W and E tiles of k_band_sk8 (m_pad x d_pad bytes
//    each, back to back in `cols`), per-SNP constants and the range flag.
enum SketchMode { kSketchRaw = 0, kSketchF32 = 1, kSketchI8 = 2 };
int sketch_plan(const ldsc_panel* panel, uint64_t d, SketchPlan** out, char* err,
                uint64_t err_len);
hipError_t launch_sketch(const SketchPlan* sp, int mode, const uint32_t* tiled, uint32_t nkb,
                         const float2* musc, uint32_t m, uint32_t m_pad, double n, double* norm,
                         void* cols, int32_t* sg, int32_t* sm, hipStream_t st);
size_t sketch_norm_bytes(const SketchPlan* sp, uint64_t m_pad);
// The per-SNP renorm scales kSketchF32 leaves after the norm shares.
const float* sketch_scales(const SketchPlan* sp, const double* norm, uint64_t m_pad);
// kSketchI8: per SNP scale, delta and A (3 x m_pad doubles), and the flag
// (non-zero when a W or E value of a real SNP left [-127, 127]).
const double* sketch_i8_consts(const SketchPlan* sp, const double* norm, uint64_t m_pad);
int32_t* sketch_i8_flag(const SketchPlan* sp, double* norm, uint64_t m_pad);
// The largest d_pad of the integer-operand form: the int32 sums of its
// products stay exact (d_pad * 127^2 < 2^31).
constexpr uint64_t kSketchI8MaxDpad = 131072;
// The one-pass form of the bucket sums (k_sketch_fused): kSfWaves waves per
// workgroup, each owning sketch_fused_bpw(nblocks) bucket blocks; a
// workgroup covers kSfWaves * bpw blocks (a range) of one 32-SNP group.
constexpr uint32_t kSfWaves = 8;
// The one-pass form is used up to this many bucket blocks (one per wave).
constexpr uint64_t kSfFusedMaxBlocks = 8;
inline uint32_t sketch_fused_bpw(uint64_t nblocks) { return nblocks <= 8 ? 1 : nblocks <= 16 ? 2 : 4; }
inline uint32_t sketch_fused_ranges(uint64_t nblocks) {
    const uint64_t per = kSfWaves * sketch_fused_bpw(nblocks);
    return (uint32_t)((nblocks + per - 1) / per);
}
// The dense one-pass form (k_sketch_dense): natural individual order, used
// up to this many bucket blocks (d <= 224).
constexpr uint64_t kSdMaxBlocks = 7;
// Bucket blocks per wave of k_sketch: the waves of a SNP pair-group split the
// d/32 bucket blocks so that the whole chip works at small d too.  Each
// wave writes one norm share slot per SNP.
constexpr uint32_t kSkChunk = 4;
// Per-SNP share slots of the norm buffer: enough for k_sketch (one per
// kSkChunk blocks), k_sketch_fused (one per wave and range) and
// k_sketch_dense (two).
inline uint32_t sketch_share_slots(uint64_t nblocks) {
    return (uint32_t)std::max<uint64_t>((nblocks + kSkChunk - 1) / kSkChunk,
                                        kSfWaves * sketch_fused_ranges(nblocks));
}
// One-pass bucket sums in the integer-operand form, straight from the
// resident .bed rows (natural order with SketchPlan::d_pcode, k_sketch_dense;
// else the chunk-local order of SketchPlan::d_local, k_sketch_fused): the W / E
// tiles, norm and A shares, range flag (as launch_sketch kSketchI8 before
// k_sketch_scale) and each SNP's integer stats in stats_out (uint4 per SNP).
// With sg_out / sm_out it writes the exact sums SG, SM instead (d x m
// column-major int32 each; norm and cols unused).
hipError_t launch_sketch_fused(const SketchPlan* sp, const ldsc_panel* panel,
                               const uint64_t* d_bed_idx, uint64_t base, uint32_t m, uint32_t m_pad,
                               uint32_t* stats_out, double* norm, void* cols, hipStream_t st,
                               int32_t* sg_out = nullptr, int32_t* sm_out = nullptr);
// After k_snp_info has the f32 mean / inverse std: per-SNP constants of the
// integer form from the shares launch_sketch_fused wrote.
hipError_t launch_sketch_fused_scale(const SketchPlan* sp, const float2* musc, uint32_t m_pad,
                                     double n, double* norm, hipStream_t st);
hipError_t launch_band_sk(const BandArgs& a, hipStream_t st);
hipError_t launch_band_sk8(const BandArgs& a, hipStream_t st);

}  // namespace ldsc
