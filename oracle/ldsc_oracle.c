/*
 * ldsc_oracle.c -- CPU restatement of the `l2` hot path of sharifhsn/ldsc.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it to compute
 * the expected answer.  It is never linked into the product library and the
 * product path never calls it.
 *
 * Each function cites the reference file:line it restates.  It follows the
 * reference's arithmetic step by step (FP64 standardized columns by default,
 * the f32 forms of `--fast-f32` and of the fused CountSketch), so that the
 * GPU engine -- which uses a different formulation (exact integer products,
 * per-SNP pair bounds, banded tiles) -- is checked against an independent
 * restatement of the reference's chunk / ring / window loop.
 *
 * The products of the chunk and window columns run as a blocked GEMM
 * (blocks of 1024 elements, 4 (f64) or 8 (f32) vector lanes each); the
 * reference uses faer's blocked GEMM, whose summation order is not
 * reproducible here, so the f64 results agree with the reference to
 * rounding (~1e-15 relative) and the f32 results to the tolerance stated in
 * the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double v4d __attribute__((vector_size(32)));
typedef float v8f __attribute__((vector_size(32)));

enum { ORC_CM = 0, ORC_KB = 1, ORC_SNP = 2 };

/* ------------------------------------------------------------------ */
/* BED header and length validation: src/bed.rs:76-109                  */
/* returns 0 ok, 1 too short, 2 bad magic, 3 not SNP-major, 4 bad length */
/* ------------------------------------------------------------------ */
int orc_bed_validate(const uint8_t* b, uint64_t len, uint64_t n_indiv, uint64_t m) {
    if (len < 3) return 1;
    if (b[0] != 0x6c || b[1] != 0x1b) return 2;
    if (b[2] != 0x01) return 3;
    uint64_t bps = (n_indiv + 3) / 4;
    if (len != 3 + bps * m) return 4;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Decode: build_lut (count_a1) [2, NaN, 1, 0] + decode_column          */
/* src/bed.rs:470-521.  Missing is reported as -1.                      */
/* keep (nullable) lists FAM rows (src/bed.rs:460-468, io.rs:53-91).    */
/* ------------------------------------------------------------------ */
static inline int code_at(const uint8_t* slab, uint64_t fam_row) {
    return (slab[fam_row >> 2] >> (2 * (fam_row & 3))) & 3;
}
static const int8_t DOSAGE_LUT[4] = {2, -1, 1, 0};

void orc_decode(const uint8_t* slabs, uint64_t bps, uint64_t n_fam, const int64_t* keep,
                uint64_t n_keep, const uint64_t* bed_idx, uint64_t m, int8_t* out) {
    uint64_t n = keep ? n_keep : n_fam;
    for (uint64_t j = 0; j < m; ++j) {
        const uint8_t* slab = slabs + bed_idx[j] * bps;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t row = keep ? (uint64_t)keep[i] : i;
            out[j * n + i] = DOSAGE_LUT[code_at(slab, row)];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Integer per-SNP stats: snp_stats_from_bytes, compute.rs:317-427, and  */
/* the het count of snp_stats.rs:24-49.  stats is m x 4 row-major:       */
/* sum, count, sum_sq, het.                                              */
/* ------------------------------------------------------------------ */
void orc_stats(const uint8_t* slabs, uint64_t bps, uint64_t n_fam, const int64_t* keep,
               uint64_t n_keep, const uint64_t* bed_idx, uint64_t m, uint32_t* stats) {
    uint64_t n = keep ? n_keep : n_fam;
    for (uint64_t j = 0; j < m; ++j) {
        const uint8_t* slab = slabs + bed_idx[j] * bps;
        uint32_t s = 0, c = 0, sq = 0, het = 0;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t row = keep ? (uint64_t)keep[i] : i;
            switch (code_at(slab, row)) {
                case 0: s += 2; c += 1; sq += 4; break;
                case 1: break;
                case 2: s += 1; c += 1; sq += 1; het += 1; break;
                default: c += 1; break;
            }
        }
        stats[4 * j + 0] = s;
        stats[4 * j + 1] = c;
        stats[4 * j + 2] = sq;
        stats[4 * j + 3] = het;
    }
}

/* ------------------------------------------------------------------ */
/* Windows: get_block_lefts_f64 + get_block_lefts_by_chr,                */
/* src/l2/window.rs:15-90.  Returns any_full_chr_window.                 */
/* ------------------------------------------------------------------ */
static void block_lefts_f64(const double* coords, uint64_t m, double max_dist, uint64_t* bl) {
    uint64_t j = 0;
    for (uint64_t i = 0; i < m; ++i) {
        while (j < m && fabs(coords[i] - coords[j]) > max_dist) j++;
        bl[i] = j;
    }
}

int orc_block_left(const uint8_t* chr, const double* cm, const uint32_t* bp, uint64_t m,
                   int mode, double value, uint64_t* bl) {
    int any_full = 0;
    uint64_t start = 0;
    double* coords = (double*)malloc(sizeof(double) * (m ? m : 1));
    uint64_t* local = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
    while (start < m) {
        uint64_t end = start + 1;
        while (end < m && chr[end] == chr[start]) end++;
        uint64_t len = end - start;
        if (mode == ORC_SNP) {
            uint64_t half = (uint64_t)value;
            if (half >= (len >= 1 ? len - 1 : 0)) any_full = 1;
            for (uint64_t i = 0; i < len; ++i) bl[start + i] = start + (i > half ? i - half : 0);
        } else {
            for (uint64_t i = 0; i < len; ++i)
                coords[i] = (mode == ORC_CM) ? cm[start + i] : (double)bp[start + i] / 1000.0;
            block_lefts_f64(coords, len, value, local);
            if (local[len - 1] == 0) any_full = 1;
            for (uint64_t i = 0; i < len; ++i) bl[start + i] = start + local[i];
        }
        start = end;
    }
    free(coords);
    free(local);
    return any_full;
}

/* ------------------------------------------------------------------ */
/* fastrand 2.3.0 (wyrand) as used by CountSketchProj::new,              */
/* compute.rs:179-192.  Restated from fastrand's published algorithm:   */
/* gen_u64 (wyrand v4.2 constants), u32(..d) via Lemire's multiply-shift */
/* with rejection, bool() = (u8)gen_u32 % 2 == 0.                        */
/* ------------------------------------------------------------------ */
static inline uint64_t wy_next(uint64_t* state) {
    uint64_t s = *state + 0x2d358dccaa6c78a5ULL;
    *state = s;
    __uint128_t t = (__uint128_t)s * (__uint128_t)(s ^ 0x8bb84b93962eacc9ULL);
    return (uint64_t)t ^ (uint64_t)(t >> 64);
}
static inline uint32_t wy_mod_u32(uint64_t* state, uint32_t n) {
    uint32_t r = (uint32_t)wy_next(state);
    uint64_t mm = (uint64_t)r * (uint64_t)n;
    uint32_t hi = (uint32_t)(mm >> 32);
    uint32_t lo = (uint32_t)mm;
    if (lo < n) {
        uint32_t t = (uint32_t)(-n) % n;
        while (lo < t) {
            r = (uint32_t)wy_next(state);
            mm = (uint64_t)r * (uint64_t)n;
            hi = (uint32_t)(mm >> 32);
            lo = (uint32_t)mm;
        }
    }
    return hi;
}
void orc_countsketch_hash(uint64_t n, uint64_t d, uint64_t seed, uint32_t* bucket,
                          int8_t* sign) {
    uint64_t st = seed;
    for (uint64_t i = 0; i < n; ++i) bucket[i] = wy_mod_u32(&st, (uint32_t)d);
    for (uint64_t i = 0; i < n; ++i) sign[i] = (((uint8_t)wy_next(&st)) % 2 == 0) ? 1 : -1;
}

/* Exact integer CountSketch buckets (the integer part of the fused
 * decode->scatter, wasm_simd.rs:1691-1784): sg[b,j] = sum s_i g_ij,
 * sm[b,j] = sum s_i [present]. */
void orc_sketch_int(const uint8_t* slabs, uint64_t bps, uint64_t n_fam, const int64_t* keep,
                    uint64_t n_keep, const uint64_t* bed_idx, uint64_t m, uint64_t d,
                    const uint32_t* bucket, const int8_t* sign, int32_t* sg, int32_t* sm) {
    uint64_t n = keep ? n_keep : n_fam;
    memset(sg, 0, sizeof(int32_t) * d * m);
    memset(sm, 0, sizeof(int32_t) * d * m);
    for (uint64_t j = 0; j < m; ++j) {
        const uint8_t* slab = slabs + bed_idx[j] * bps;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t row = keep ? (uint64_t)keep[i] : i;
            int g = DOSAGE_LUT[code_at(slab, row)];
            if (g < 0) continue;
            sg[j * d + bucket[i]] += sign[i] * g;
            sm[j * d + bucket[i]] += sign[i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Standardization: normalize_col_f64_with_stats / _f32_with_stats       */
/* (src/l2/normalize.rs:71-154) and the fused-sketch stats pass          */
/* (compute.rs:629-668).                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
    double avg, inv_std, maf;
    float avg_f32, inv_f32;
} colstat;

static colstat make_colstat(uint32_t sum, uint32_t count, uint32_t sum_sq, uint64_t n) {
    colstat cs;
    double avg = count > 0 ? (double)sum / (double)count : 0.0;
    double freq = avg / 2.0;
    if (freq < 0.0) freq = 0.0;
    if (freq > 1.0) freq = 1.0;
    cs.maf = freq < 1.0 - freq ? freq : 1.0 - freq;
    double css = (double)sum_sq - (double)count * avg * avg;
    double var = css / (double)n;
    double sd = sqrt(var);
    cs.avg = avg;
    cs.inv_std = sd > 0.0 ? 1.0 / sd : 0.0;
    cs.avg_f32 = (float)avg;
    cs.inv_f32 = sd > 0.0 ? (float)(1.0 / sd) : 0.0f;
    return cs;
}

/* f32 standardized columns in the --fast-f32 form (normalize.rs:115-154). */
void orc_std_f32(const uint8_t* slabs, uint64_t bps, uint64_t n_fam, const int64_t* keep,
                 uint64_t n_keep, const uint64_t* bed_idx, uint64_t m, float* out) {
    uint64_t n = keep ? n_keep : n_fam;
    for (uint64_t j = 0; j < m; ++j) {
        uint32_t st[4];
        orc_stats(slabs, bps, n_fam, keep, n_keep, bed_idx + j, 1, st);
        colstat cs = make_colstat(st[0], st[1], st[2], n);
        const uint8_t* slab = slabs + bed_idx[j] * bps;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t row = keep ? (uint64_t)keep[i] : i;
            int g = DOSAGE_LUT[code_at(slab, row)];
            out[j * n + i] = g < 0 ? 0.0f : ((float)g - cs.avg_f32) * cs.inv_f32;
        }
    }
}

/* per_snp_kappa, compute.rs:525-538 (from integer raw moments). */
static double per_snp_kappa(uint32_t sum, uint32_t count, uint32_t sum_sq, uint32_t sum_cube,
                            uint32_t sum_4th) {
    if (count == 0) return 0.0;
    double cf = (double)count;
    double mean = (double)sum / cf;
    double mean2 = mean * mean; /* mean.powi(4): (x*x)*(x*x) */
    double m4 = (double)sum_4th - 4.0 * mean * (double)sum_cube +
                6.0 * mean * mean * (double)sum_sq - 3.0 * (mean2 * mean2) * cf;
    double css = (double)sum_sq - cf * mean * mean;
    if (css <= 0.0) return 0.0;
    return m4 / (css * css);
}

/* quadratic_sketch_correction[_maf_aware], compute.rs:947-998. */
static inline double quad_corr(double val, double n_inv_sq, double d_f, double r2u_a,
                               double r2u_b, double kj, double kk) {
    double t = val * val * n_inv_sq;
    double dm3 = d_f - 3.0 - 2.0 * (kj + kk);
    double disc = dm3 * dm3 + 8.0 * (d_f * t - 1.0);
    double rs = disc >= 0.0 ? (-dm3 + sqrt(disc)) * 0.25 : 0.0;
    return rs * r2u_a + r2u_b;
}

/* ------------------------------------------------------------------ */
/* Blocked products of many columns (the faer GEMMs of                    */
/* compute.rs:1701-1708, 1802-1809, 2010-2017, 2171-2178):               */
/* G[j * na + i] = <A_i, B_j> for columns of length n given by pointers. */
/* The length is cut into blocks of ORC_KB elements; within a block a    */
/* 4 x 2 register tile of vector accumulators runs over it and its lane  */
/* sums are added to G, so every product sums in the same fixed order    */
/* whatever the threading.  Threads take (group of B, range of A) tasks. */
/* ------------------------------------------------------------------ */
#define ORC_KB 1024
#define ORC_JG 32
#define ORC_IR 64

static void gram_f64(const double* const* A, uint64_t na, const double* const* B, uint64_t nb,
                     uint64_t n, double* G) {
    memset(G, 0, sizeof(double) * na * nb);
    const uint64_t ngroups = (nb + ORC_JG - 1) / ORC_JG, nranges = (na + ORC_IR - 1) / ORC_IR;
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t task = 0; task < ngroups * nranges; ++task) {
        const uint64_t g = task / nranges, h = task % nranges;
        const uint64_t j0 = g * ORC_JG, j1 = j0 + ORC_JG < nb ? j0 + ORC_JG : nb;
        const uint64_t i0 = h * ORC_IR, i1 = i0 + ORC_IR < na ? i0 + ORC_IR : na;
        for (uint64_t k0 = 0; k0 < n; k0 += ORC_KB) {
            const uint64_t k1 = k0 + ORC_KB < n ? k0 + ORC_KB : n;
            const uint64_t kv = k0 + ((k1 - k0) / 4) * 4;  /* vector part [k0, kv) */
            for (uint64_t i = i0; i < i1; i += 4) {
                const uint64_t ni = i1 - i < 4 ? i1 - i : 4;
                for (uint64_t j = j0; j < j1; j += 2) {
                    const uint64_t nj = j1 - j < 2 ? j1 - j : 2;
                    v4d acc[4][2];
                    memset(acc, 0, sizeof(acc));
                    /* edge tiles repeat a valid column; its sums are not stored */
                    const double* pa[4];
                    const double* pb[2];
                    for (uint64_t u = 0; u < 4; ++u) pa[u] = A[i + (u < ni ? u : 0)];
                    for (uint64_t v = 0; v < 2; ++v) pb[v] = B[j + (v < nj ? v : 0)];
                    for (uint64_t k = k0; k < kv; k += 4) {
                        v4d a0, a1, a2, a3, b0, b1;
                        memcpy(&a0, pa[0] + k, 32);
                        memcpy(&a1, pa[1] + k, 32);
                        memcpy(&a2, pa[2] + k, 32);
                        memcpy(&a3, pa[3] + k, 32);
                        memcpy(&b0, pb[0] + k, 32);
                        memcpy(&b1, pb[1] + k, 32);
                        acc[0][0] += a0 * b0; acc[0][1] += a0 * b1;
                        acc[1][0] += a1 * b0; acc[1][1] += a1 * b1;
                        acc[2][0] += a2 * b0; acc[2][1] += a2 * b1;
                        acc[3][0] += a3 * b0; acc[3][1] += a3 * b1;
                    }
                    for (uint64_t u = 0; u < ni; ++u)
                        for (uint64_t v = 0; v < nj; ++v) {
                            double s = (acc[u][v][0] + acc[u][v][1]) +
                                       (acc[u][v][2] + acc[u][v][3]);
                            for (uint64_t k = kv; k < k1; ++k) s += A[i + u][k] * B[j + v][k];
                            G[(j + v) * na + i + u] += s;
                        }
                }
            }
        }
    }
}

static void gram_f32(const float* const* A, uint64_t na, const float* const* B, uint64_t nb,
                     uint64_t n, float* G) {
    memset(G, 0, sizeof(float) * na * nb);
    const uint64_t ngroups = (nb + ORC_JG - 1) / ORC_JG, nranges = (na + ORC_IR - 1) / ORC_IR;
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t task = 0; task < ngroups * nranges; ++task) {
        const uint64_t g = task / nranges, h = task % nranges;
        const uint64_t j0 = g * ORC_JG, j1 = j0 + ORC_JG < nb ? j0 + ORC_JG : nb;
        const uint64_t i0 = h * ORC_IR, i1 = i0 + ORC_IR < na ? i0 + ORC_IR : na;
        for (uint64_t k0 = 0; k0 < n; k0 += ORC_KB) {
            const uint64_t k1 = k0 + ORC_KB < n ? k0 + ORC_KB : n;
            const uint64_t kv = k0 + ((k1 - k0) / 8) * 8;
            for (uint64_t i = i0; i < i1; i += 4) {
                const uint64_t ni = i1 - i < 4 ? i1 - i : 4;
                for (uint64_t j = j0; j < j1; j += 2) {
                    const uint64_t nj = j1 - j < 2 ? j1 - j : 2;
                    v8f acc[4][2];
                    memset(acc, 0, sizeof(acc));
                    const float* pa[4];
                    const float* pb[2];
                    for (uint64_t u = 0; u < 4; ++u) pa[u] = A[i + (u < ni ? u : 0)];
                    for (uint64_t v = 0; v < 2; ++v) pb[v] = B[j + (v < nj ? v : 0)];
                    for (uint64_t k = k0; k < kv; k += 8) {
                        v8f a0, a1, a2, a3, b0, b1;
                        memcpy(&a0, pa[0] + k, 32);
                        memcpy(&a1, pa[1] + k, 32);
                        memcpy(&a2, pa[2] + k, 32);
                        memcpy(&a3, pa[3] + k, 32);
                        memcpy(&b0, pb[0] + k, 32);
                        memcpy(&b1, pb[1] + k, 32);
                        acc[0][0] += a0 * b0; acc[0][1] += a0 * b1;
                        acc[1][0] += a1 * b0; acc[1][1] += a1 * b1;
                        acc[2][0] += a2 * b0; acc[2][1] += a2 * b1;
                        acc[3][0] += a3 * b0; acc[3][1] += a3 * b1;
                    }
                    for (uint64_t u = 0; u < ni; ++u)
                        for (uint64_t v = 0; v < nj; ++v) {
                            const v8f x = acc[u][v];
                            float s = ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
                            for (uint64_t k = kv; k < k1; ++k) s += A[i + u][k] * B[j + v][k];
                            G[(j + v) * na + i + u] += s;
                        }
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* compute_ldscore_global, src/l2/compute.rs:1009-2447.                  */
/*                                                                        */
/* slabs: the .bed body (header stripped), bps bytes per SNP row.        */
/* annot: m x K column-major (NULL = all ones, K = 1).                   */
/* sketch_d: 0 = exact.  use_f32: --fast-f32 (forced on by run() when    */
/* sketching, mod.rs:393).  Returns 0 ok, 1 chunk_c == 0,                */
/* 2 sketch d out of range.  l2 is m x K column-major, maf is m.         */
/* ------------------------------------------------------------------ */
/* Test hook: a caller-given projection (bucket, sign over the n individuals
 * in panel order, sketch dimension d) replaces CountSketchProj::new(n, d, 42)
 * in orc_l2 while it is set, and d = n is then allowed (the reference's
 * range check, compute.rs:1083-1088, stays for the hashed projection).  With
 * an identity projection (d = n, bucket i = i, all signs +1) the sketched
 * column is the standardized column itself, so the sketch path's chain
 * (scatter, renorm, product, quadratic inversion) can be checked apart from
 * the hash.  NULL arrays clear the hook. */
static const uint32_t* g_proj_bucket = NULL;
static const int8_t* g_proj_sign = NULL;
static uint64_t g_proj_n = 0, g_proj_d = 0;
void orc_set_projection(const uint32_t* bucket, const int8_t* sign, uint64_t n, uint64_t d) {
    g_proj_bucket = bucket && sign ? bucket : NULL;
    g_proj_sign = bucket && sign ? sign : NULL;
    g_proj_n = n;
    g_proj_d = d;
}

/* per_snp_kappa of each listed SNP from the integer moments the sketch
 * path counts (compute.rs:525-538, moments as in compute.rs:429-523). */
void orc_kappa(const uint8_t* slabs, uint64_t bps, uint64_t n_fam, const int64_t* keep,
               uint64_t n_keep, const uint64_t* bed_idx, uint64_t m, double* out) {
    const uint64_t n = keep ? n_keep : n_fam;
    for (uint64_t j = 0; j < m; ++j) {
        uint32_t st[4];
        orc_stats(slabs, bps, n_fam, keep, n_keep, bed_idx + j, 1, st);
        const uint8_t* slab = slabs + bed_idx[j] * bps;
        uint32_t cube = 0, fourth = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t row = keep ? (uint64_t)keep[i] : i;
            const int gg = DOSAGE_LUT[code_at(slab, row)];
            if (gg > 0) {
                cube += (uint32_t)(gg * gg * gg);
                fourth += (uint32_t)(gg * gg * gg * gg);
            }
        }
        out[j] = per_snp_kappa(st[0], st[1], st[2], cube, fourth);
    }
}

int orc_l2(const uint8_t* slabs, uint64_t bps, uint64_t n_fam, const int64_t* keep,
           uint64_t n_keep, const uint8_t* chr, const double* cm, const uint32_t* bp,
           const uint64_t* bed_idx, uint64_t m, int mode, double wval, uint64_t chunk_c,
           int use_f32, uint64_t sketch_d, int maf_aware, int masking, double pq_exp,
           const double* annot, uint64_t K, double* l2, double* maf_out, int nthreads) {
    if (chunk_c < 1) return 1; /* compute.rs:1035-1039 */
    if (m == 0) return 0;
    const uint64_t n = keep ? n_keep : n_fam;
    const int given_proj = sketch_d && g_proj_bucket && g_proj_n == n && g_proj_d == sketch_d;
    if (sketch_d && !given_proj && !(sketch_d >= 3 && sketch_d < n)) return 2; /* :1083-1088 */
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    uint64_t* bl = (uint64_t*)malloc(sizeof(uint64_t) * m);
    orc_block_left(chr, cm, bp, m, mode, wval, bl);

    /* r2_unbiased constants, compute.rs:1067-1077 */
    const double nd = (double)n, n_inv = 1.0 / nd;
    const double r2u_denom = n > 2 ? nd - 2.0 : nd;
    const double r2u_a = 1.0 + 1.0 / r2u_denom, r2u_b = -1.0 / r2u_denom;
    const double n_inv_sq_r2u_a = n_inv * n_inv * r2u_a;
    const double n_inv_sq = n_inv * n_inv;
    const uint64_t gemm_n = sketch_d ? sketch_d : n;

    /* CountSketchProj::new(n, d, 42), compute.rs:1147 */
    uint32_t* bucket = NULL;
    int8_t* sgn = NULL;
    if (sketch_d) {
        bucket = (uint32_t*)malloc(sizeof(uint32_t) * n);
        sgn = (int8_t*)malloc(n);
        if (given_proj) {
            memcpy(bucket, g_proj_bucket, sizeof(uint32_t) * n);
            memcpy(sgn, g_proj_sign, n);
        } else {
            orc_countsketch_hash(n, sketch_d, 42, bucket, sgn);
        }
    }

    /* block sizes and ring, compute.rs:1164-1181 */
    uint64_t* bsz = (uint64_t*)malloc(sizeof(uint64_t) * m);
    uint64_t max_w = 0;
    for (uint64_t i = 0; i < m; ++i) {
        uint64_t dist = i - bl[i];
        bsz[i] = dist == 0 ? 0 : ((dist + chunk_c - 1) / chunk_c) * chunk_c;
        if (bsz[i] > max_w) max_w = bsz[i];
    }
    uint64_t ring = max_w + 8 * chunk_c;
    if (ring < 1) ring = 1;
    ring = ((ring + chunk_c - 1) / chunk_c) * chunk_c;
    double* ring64 = use_f32 ? NULL : (double*)calloc(gemm_n * ring, sizeof(double));
    float* ring32 = use_f32 ? (float*)calloc(gemm_n * ring, sizeof(float)) : NULL;

    double* kurt = (double*)calloc(m, sizeof(double));
    double* pq = (double*)malloc(sizeof(double) * m);
    for (uint64_t i = 0; i < m; ++i) pq[i] = 1.0;
    memset(l2, 0, sizeof(double) * m * K);

    /* A_eff accessor: annot (or ones) times pq when pq_exp is set */
    const int has_pq = !isnan(pq_exp);

    double* r2u_bb = (double*)malloc(sizeof(double) * chunk_c * chunk_c);
    double* r2u_ab = (double*)malloc(sizeof(double) * (max_w ? max_w : 1) * chunk_c);
    /* column pointers and products of the blocked GEMMs */
    const uint64_t gram_len = ((max_w > chunk_c ? max_w : chunk_c) + 1) * chunk_c;
    const double** colp64 = (const double**)malloc(sizeof(double*) * chunk_c);
    const double** winp64 = (const double**)malloc(sizeof(double*) * (max_w + 1));
    const float** colp32 = (const float**)malloc(sizeof(float*) * chunk_c);
    const float** winp32 = (const float**)malloc(sizeof(float*) * (max_w + 1));
    double* gram64 = use_f32 ? NULL : (double*)malloc(sizeof(double) * gram_len);
    float* gram32 = use_f32 ? (float*)malloc(sizeof(float) * gram_len) : NULL;
    double* aeff_chunk = (double*)malloc(sizeof(double) * chunk_c * K);
    double* aeff_win = (double*)malloc(sizeof(double) * (max_w ? max_w : 1) * K);
    uint64_t win_lo = 0; /* window = [win_lo, chunk_start) (ring contents) */
    uint8_t* rawcol = (uint8_t*)malloc(n);

    for (uint64_t cs0 = 0; cs0 < m; cs0 += chunk_c) {
        const uint64_t ce = cs0 + chunk_c < m ? cs0 + chunk_c : m;
        const uint64_t c = ce - cs0;
        uint64_t b = bsz[cs0];
        if (b > cs0) b = cs0;
        const uint64_t a_left = cs0 - b; /* :1288-1299 */
        if (win_lo < a_left) win_lo = a_left;
        const int contiguous = bed_idx[ce - 1] == bed_idx[cs0] + c - 1; /* :1307-1318 */

        /* decode + normalize (+ sketch) each chunk column into its ring slot */
        for (uint64_t j = 0; j < c; ++j) {
            const uint64_t g = cs0 + j;
            const uint64_t slot = g % ring;
            uint32_t st[4];
            orc_stats(slabs, bps, n_fam, keep, n_keep, bed_idx + g, 1, st);
            colstat csx = make_colstat(st[0], st[1], st[2], n);
            const uint8_t* slab = slabs + bed_idx[g] * bps;
            for (uint64_t i = 0; i < n; ++i) {
                uint64_t row = keep ? (uint64_t)keep[i] : i;
                rawcol[i] = (uint8_t)code_at(slab, row);
            }
            if (sketch_d && maf_aware) {
                uint32_t cube = 0, fourth = 0;
                for (uint64_t i = 0; i < n; ++i) {
                    int gg = DOSAGE_LUT[rawcol[i]];
                    if (gg > 0) {
                        cube += (uint32_t)(gg * gg * gg);
                        fourth += (uint32_t)(gg * gg * gg * gg);
                    }
                }
                kurt[g] = per_snp_kappa(st[0], st[1], st[2], cube, fourth);
            }
            double maf_v;
            if (!sketch_d) {
                if (use_f32) {
                    float* col = ring32 + slot * gemm_n;
                    for (uint64_t i = 0; i < n; ++i) {
                        int gg = DOSAGE_LUT[rawcol[i]];
                        col[i] = gg < 0 ? 0.0f : ((float)gg - csx.avg_f32) * csx.inv_f32;
                    }
                    maf_v = (double)(float)csx.maf; /* normalize.rs:153 */
                } else {
                    double* col = ring64 + slot * gemm_n;
                    for (uint64_t i = 0; i < n; ++i) {
                        int gg = DOSAGE_LUT[rawcol[i]];
                        col[i] = gg < 0 ? 0.0 : ((double)gg - csx.avg) * csx.inv_std;
                    }
                    maf_v = csx.maf;
                }
            } else if (use_f32) {
                /* fused f32 scatter (contiguous chunk) or normalize+project_f32:
                 * identical values; MAF is f64 on the fused path (compute.rs:639)
                 * and f32-rounded on the unfused one (compute.rs:1509-1511). */
                float* col = ring32 + slot * gemm_n;
                const float vals[4] = {(2.0f - csx.avg_f32) * csx.inv_f32, 0.0f,
                                       (1.0f - csx.avg_f32) * csx.inv_f32,
                                       -csx.avg_f32 * csx.inv_f32};
                memset(col, 0, sizeof(float) * gemm_n);
                for (uint64_t i = 0; i < n; ++i) {
                    float v;
                    if (contiguous) {
                        v = vals[rawcol[i]];
                    } else {
                        int gg = DOSAGE_LUT[rawcol[i]];
                        v = gg < 0 ? 0.0f : ((float)gg - csx.avg_f32) * csx.inv_f32;
                    }
                    col[bucket[i]] += (float)sgn[i] * v;
                }
                double nrm = 0.0;
                for (uint64_t q = 0; q < gemm_n; ++q) nrm += (double)col[q] * (double)col[q];
                if (nrm > 0.0) {
                    float sc = (float)sqrt(nd / nrm);
                    for (uint64_t q = 0; q < gemm_n; ++q) col[q] *= sc;
                }
                maf_v = contiguous ? csx.maf : (double)(float)csx.maf;
            } else {
                /* f64 sketch (fused f64 / project_f64) */
                double* col = ring64 + slot * gemm_n;
                const double vals[4] = {(2.0 - csx.avg) * csx.inv_std, 0.0,
                                        (1.0 - csx.avg) * csx.inv_std, -csx.avg * csx.inv_std};
                memset(col, 0, sizeof(double) * gemm_n);
                for (uint64_t i = 0; i < n; ++i) {
                    double v;
                    if (contiguous) {
                        v = vals[rawcol[i]];
                    } else {
                        int gg = DOSAGE_LUT[rawcol[i]];
                        v = gg < 0 ? 0.0 : ((double)gg - csx.avg) * csx.inv_std;
                    }
                    col[bucket[i]] += (double)sgn[i] * v;
                }
                double nrm = 0.0;
                for (uint64_t q = 0; q < gemm_n; ++q) nrm += col[q] * col[q];
                if (nrm > 0.0) {
                    double sc = sqrt(nd / nrm);
                    for (uint64_t q = 0; q < gemm_n; ++q) col[q] *= sc;
                }
                maf_v = csx.maf;
            }
            maf_out[g] = maf_v;
            if (has_pq) pq[g] = pow(maf_v * (1.0 - maf_v), pq_exp);
        }

        /* A_eff for the chunk and the window (compute.rs:1595-1631, 1913-1947) */
        int chunk_all_zero = 1, window_all_zero = 1;
        for (uint64_t k = 0; k < K; ++k)
            for (uint64_t j = 0; j < c; ++j) {
                double a = annot ? annot[k * m + cs0 + j] : 1.0;
                if (has_pq) a *= pq[cs0 + j];
                aeff_chunk[k * c + j] = a;
                if (a != 0.0) chunk_all_zero = 0;
            }
        const uint64_t w = cs0 - win_lo;
        for (uint64_t k = 0; k < K; ++k)
            for (uint64_t i = 0; i < w; ++i) {
                double a = annot ? annot[k * m + win_lo + i] : 1.0;
                if (has_pq) a *= pq[win_lo + i];
                aeff_win[k * w + i] = a;
                if (a != 0.0) window_all_zero = 0;
            }

        /* ---- B x B (compute.rs:1633-1893) ---- */
        if (!chunk_all_zero) {
            for (uint64_t j = 0; j < c; ++j) {
                const uint64_t sj = (cs0 + j) % ring;
                if (use_f32) colp32[j] = ring32 + sj * gemm_n;
                else colp64[j] = ring64 + sj * gemm_n;
            }
            if (use_f32) gram_f32(colp32, c, colp32, c, gemm_n, gram32);
            else gram_f64(colp64, c, colp64, c, gemm_n, gram64);
            for (uint64_t j = 0; j < c; ++j) {
                r2u_bb[j * c + j] = 1.0;
                for (uint64_t k = 0; k < j; ++k) {
                    double val = use_f32 ? (double)gram32[j * c + k] : gram64[j * c + k];
                    double r;
                    if (sketch_d)
                        r = quad_corr(val, n_inv_sq, (double)sketch_d, r2u_a, r2u_b,
                                      maf_aware ? kurt[cs0 + j] : 0.0,
                                      maf_aware ? kurt[cs0 + k] : 0.0);
                    else
                        r = val * val * n_inv_sq_r2u_a + r2u_b;
                    r2u_bb[j * c + k] = r;
                    r2u_bb[k * c + j] = r;
                }
            }
            if (masking) { /* compute.rs:1855-1867 */
                for (uint64_t j = 1; j < c; ++j) {
                    uint64_t bl_j = bl[cs0 + j];
                    if (bl_j <= cs0) continue;
                    uint64_t cut = bl_j - cs0;
                    for (uint64_t k = 0; k < (cut < j ? cut : j); ++k) {
                        r2u_bb[j * c + k] = 0.0;
                        r2u_bb[k * c + j] = 0.0;
                    }
                }
            }
            /* l2[chunk] += r2u_bb . A_eff[chunk] (compute.rs:1869-1893) */
            for (uint64_t kk = 0; kk < K; ++kk)
                for (uint64_t j = 0; j < c; ++j) {
                    double s = 0.0;
                    for (uint64_t k = 0; k < c; ++k)
                        s += r2u_bb[k * c + j] * aeff_chunk[kk * c + k];
                    l2[kk * m + cs0 + j] += s;
                }
        }

        /* ---- A x B (compute.rs:1897-2315) ---- */
        if (w > 0 && !(chunk_all_zero && window_all_zero)) {
            for (uint64_t j = 0; j < c; ++j) {
                const uint64_t sj = (cs0 + j) % ring;
                if (use_f32) colp32[j] = ring32 + sj * gemm_n;
                else colp64[j] = ring64 + sj * gemm_n;
            }
            for (uint64_t i = 0; i < w; ++i) {
                const uint64_t si = (win_lo + i) % ring;
                if (use_f32) winp32[i] = ring32 + si * gemm_n;
                else winp64[i] = ring64 + si * gemm_n;
            }
            if (use_f32) gram_f32(winp32, w, colp32, c, gemm_n, gram32);
            else gram_f64(winp64, w, colp64, c, gemm_n, gram64);
            for (uint64_t j = 0; j < c; ++j) {
                for (uint64_t i = 0; i < w; ++i) {
                    double val = use_f32 ? (double)gram32[j * w + i] : gram64[j * w + i];
                    double r;
                    if (sketch_d)
                        r = quad_corr(val, n_inv_sq, (double)sketch_d, r2u_a, r2u_b,
                                      maf_aware ? kurt[win_lo + i] : 0.0,
                                      maf_aware ? kurt[cs0 + j] : 0.0);
                    else
                        r = val * val * n_inv_sq_r2u_a + r2u_b;
                    r2u_ab[j * w + i] = r;
                }
            }
            if (masking) { /* compute.rs:2244-2257 */
                uint64_t cutoff = 0;
                for (uint64_t j = 0; j < c; ++j) {
                    uint64_t bl_j = bl[cs0 + j];
                    while (cutoff < w && win_lo + cutoff < bl_j) cutoff++;
                    for (uint64_t i = 0; i < cutoff; ++i) r2u_ab[j * w + i] = 0.0;
                }
            }
            if (!chunk_all_zero) /* l2[window] += r2u_ab . A_eff[chunk] */
                for (uint64_t kk = 0; kk < K; ++kk)
                    for (uint64_t i = 0; i < w; ++i) {
                        double s = 0.0;
                        for (uint64_t j = 0; j < c; ++j)
                            s += r2u_ab[j * w + i] * aeff_chunk[kk * c + j];
                        l2[kk * m + win_lo + i] += s;
                    }
            if (!window_all_zero) /* l2[chunk] += r2u_ab^T . A_eff[window] */
                for (uint64_t kk = 0; kk < K; ++kk)
                    for (uint64_t j = 0; j < c; ++j) {
                        double s = 0.0;
                        for (uint64_t i = 0; i < w; ++i)
                            s += r2u_ab[j * w + i] * aeff_win[kk * w + i];
                        l2[kk * m + cs0 + j] += s;
                    }
        }
        /* ring store: columns already live in slot g % ring (compute.rs:2320-2377) */
    }

    free(bl); free(bsz); free(ring64); free(ring32); free(kurt); free(pq);
    free(r2u_bb); free(r2u_ab); free(aeff_chunk); free(aeff_win); free(rawcol);
    free(colp64); free(winp64); free(colp32); free(winp32); free(gram64); free(gram32);
    free(bucket); free(sgn);
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* scatter_one_column_f32_scalar, all-iids path (wasm_simd.rs:1691-1784):
 * byte LUT of normalized values [(2-mean)inv, 0, (1-mean)inv, -mean*inv],
 * sequential scatter-add in individual order, f64 norm, f32 rescale. */
void orc_scatter_col_f32(const uint8_t* slab, uint64_t n, const uint32_t* bucket,
                         const float* sign, float mean, float inv, double n_norm, uint64_t d,
                         float* out) {
    const float vals[4] = {(2.0f - mean) * inv, 0.0f, (1.0f - mean) * inv, -mean * inv};
    memset(out, 0, sizeof(float) * d);
    for (uint64_t i = 0; i < n; ++i) out[bucket[i]] += sign[i] * vals[code_at(slab, i)];
    double nrm = 0.0;
    for (uint64_t q = 0; q < d; ++q) nrm += (double)out[q] * (double)out[q];
    if (nrm > 0.0) {
        float sc = (float)sqrt(n_norm / nrm);
        for (uint64_t q = 0; q < d; ++q) out[q] *= sc;
    }
}

double orc_quad_corr(double val, double n_inv_sq, double d_f, double r2u_a, double r2u_b,
                     double kj, double kk) {
    return quad_corr(val, n_inv_sq, d_f, r2u_a, r2u_b, kj, kk);
}
