// band_common.h -- device code shared by the exact and the sketch band
// kernels: the tile list, the XCD-aware tile mapping, the masked r2u
// epilogue and the deterministic per-tile annotation contraction.
#pragma once

#include "device.h"

namespace ldsc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------
// Kernel 3: banded contraction
// ------------------------------------------------------------------------
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
    const double* kappa;   // per-SNP kurtosis of the MAF-aware sketch or nullptr
    uint32_t nkb, ntiles, per_xcd, m, m_pad, K;
    double r2u_a, r2u_b, n_inv_sq_r2u_a, n_inv_sq, d_f;
    double n_pad;          // individuals of the tiles, padding included
};

// Tile -> XCD mapping: workgroups are dealt to the 8 XCDs round-robin, so
// workgroup b runs on XCD b%8.  Give each XCD one contiguous stretch of the
// (I, J)-sorted tile list: neighbouring tiles share SNP groups, which then
// hit in that XCD's L2.
__device__ inline int tile_of_block(const BandArgs& a) {
    const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const uint32_t t = xcd * a.per_xcd + slot;
    return t < a.ntiles ? (int)t : -1;
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

struct EpiShared {
    double R[32][65];     // one 32-row half of the tile's masked r2u
    double ar[64][8];     // A_eff of the tile's row SNPs (chunk of 8 annotations)
    double ac[64][8];     // A_eff of the tile's column SNPs
    int2 csr[64], csc[64];
    double rsr[64], rsc[64];
    uint32_t lor[64];
};

__device__ __forceinline__ int row_of(int bi, int r, int lane) {
    return 32 * bi + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// r2u (masked, 64 per lane in MFMA output order) -> per-tile partial sums
// for annotations kc0 .. kc0+kcn-1 (kcn <= KC; KC = 1 lets the common
// single-annotation case run without per-annotation predicates).
//   rowp[tile][k][a] = sum_b R[a][b] * A_eff[col b][k]
//   colp[tile][k][b] = sum_a R[a][b] * A_eff[row a][k]
template <int KC>
__device__ __forceinline__ void band_contract_chunk(const BandArgs& a, EpiShared& sh, int tile,
                                                    uint32_t I, uint32_t J,
                                                    const double (&r2u)[2][2][16], uint32_t kc0,
                                                    uint32_t kcn_in) {
    const int lane = threadIdx.x;
    const uint32_t kcn = KC == 1 ? 1u : kcn_in;
    __syncthreads();
    for (uint32_t kk = 0; kk < kcn; ++kk) {
        sh.ar[lane][kk] = a.aeff[(size_t)(kc0 + kk) * a.m_pad + (size_t)I * kTile + lane];
        sh.ac[lane][kk] = a.aeff[(size_t)(kc0 + kk) * a.m_pad + (size_t)J * kTile + lane];
    }
    double cacc[KC];
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) cacc[kk] = 0.0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        __syncthreads();
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                sh.R[row_of(0, r, lane)][32 * bj + (lane & 31)] = r2u[bi][bj][r];
        __syncthreads();
        // row role: lane covers row lane%32 over columns of its half
        double racc[KC];
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) racc[kk] = 0.0;
        const int rr = lane & 31, b0 = 32 * (lane >> 5);
        for (int b = 0; b < 32; ++b) {
            const double rv = sh.R[rr][b0 + b];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
                if ((uint32_t)kk < kcn) racc[kk] += rv * sh.ac[b0 + b][kk];
        }
#pragma unroll
        for (int kk = 0; kk < KC; ++kk)
            if ((uint32_t)kk < kcn) racc[kk] += __shfl_xor(racc[kk], 32);
        if (lane < 32)
            for (uint32_t kk = 0; kk < kcn; ++kk)
                a.rowp[((size_t)tile * a.K + kc0 + kk) * kTile + 32 * bi + lane] = racc[kk];
        // column role: lane = column, rows of this half
        for (int r = 0; r < 32; ++r) {
            const double rv = sh.R[r][lane];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
                if ((uint32_t)kk < kcn) cacc[kk] += rv * sh.ar[32 * bi + r][kk];
        }
    }
    for (uint32_t kk = 0; kk < kcn; ++kk)
        a.colp[((size_t)tile * a.K + kc0 + kk) * kTile + lane] = cacc[kk];
}

__device__ __forceinline__ void band_contract(const BandArgs& a, EpiShared& sh, int tile,
                                              uint32_t I, uint32_t J,
                                              const double (&r2u)[2][2][16]) {
    if (a.K == 1) {
        band_contract_chunk<1>(a, sh, tile, I, J, r2u, 0, 1);
        return;
    }
    for (uint32_t kc0 = 0; kc0 < a.K; kc0 += 8)
        band_contract_chunk<8>(a, sh, tile, I, J, r2u, kc0, a.K - kc0 < 8 ? a.K - kc0 : 8);
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


inline void plan_band(const std::vector<uint64_t>& lo, uint64_t m, BandPlan& p) {
    p.m_pad = (uint32_t)round_up(m, kTile);
    p.nb = p.m_pad / kTile;
    p.j0.assign(p.nb, 0);
    p.tstart.assign(p.nb + 1, 0);
    p.tiles.clear();
    for (uint32_t I = 0; I < p.nb; ++I) {
        const uint64_t first = (uint64_t)I * kTile;
        const int J0 = (int)(lo[first] / kTile);
        p.j0[I] = J0;
        p.tstart[I + 1] = p.tstart[I] + ((int)I - J0 + 1);
        for (int J = J0; J <= (int)I; ++J) p.tiles.push_back(make_int2((int)I, J));
    }
    p.ntiles = (uint32_t)p.tiles.size();
}

// CountSketch (sketch.hip)
int sketch_plan(const ldsc_panel* panel, uint64_t d, SketchPlan** out, char* err,
                uint64_t err_len);
hipError_t launch_sketch(const SketchPlan* sp, const uint32_t* tiled, uint32_t nkb,
                         const float2* musc, uint32_t m, uint32_t m_pad, double n, double* norm,
                         float* sk, int32_t* sg, int32_t* sm, hipStream_t st);
size_t sketch_norm_bytes(const SketchPlan* sp, uint64_t m_pad);
hipError_t launch_band_sk(const BandArgs& a, hipStream_t st);

}  // namespace ldsc
