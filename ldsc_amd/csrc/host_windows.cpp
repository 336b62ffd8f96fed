// host_windows.cpp -- host-side index work of the l2 path: LD windows, the
// per-SNP pair bound that encodes the reference's chunk/ring/mask inclusion
// rule, and the seeded CountSketch hash.
#include <cmath>
#include <cstring>

#include "internal.h"

namespace ldsc {

// get_block_lefts_f64 (src/l2/window.rs:15-28): first j with |c_i - c_j| <= W,
// advancing while the distance is strictly greater.
static void block_lefts_f64(const double* coords, uint64_t m, double max_dist,
                            uint64_t* out) {
    uint64_t j = 0;
    for (uint64_t i = 0; i < m; ++i) {
        while (j < m && std::fabs(coords[i] - coords[j]) > max_dist) ++j;
        out[i] = j;
    }
}

bool block_lefts(const ldsc_snp* snps, uint64_t m, int mode, double value,
                 std::vector<uint64_t>& bl) {
    bl.assign(m, 0);
    bool any_full = false;
    std::vector<double> coords;
    std::vector<uint64_t> local;
    uint64_t start = 0;
    while (start < m) {  // contiguous runs of one chromosome code (window.rs:41-49)
        uint64_t end = start + 1;
        while (end < m && snps[end].chr == snps[start].chr) ++end;
        const uint64_t len = end - start;
        if (mode == LDSC_WINDOW_SNP) {
            const uint64_t half = (uint64_t)value;
            if (half >= (len > 0 ? len - 1 : 0)) any_full = true;
            for (uint64_t i = 0; i < len; ++i) bl[start + i] = start + (i > half ? i - half : 0);
        } else {
            coords.resize(len);
            local.resize(len);
            for (uint64_t i = 0; i < len; ++i)
                coords[i] = mode == LDSC_WINDOW_CM ? snps[start + i].cm
                                                   : (double)snps[start + i].bp / 1000.0;
            block_lefts_f64(coords.data(), len, value, local.data());
            if (local[len - 1] == 0) any_full = true;
            for (uint64_t i = 0; i < len; ++i) bl[start + i] = start + local[i];
        }
        start = end;
    }
    return any_full;
}

int check_config(const ldsc_l2_config* cfg, char* err, uint64_t err_len) {
    if (!cfg) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "config is NULL");
    if (cfg->chunk_size < 1)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG,
                    "compute_ldscore_global: chunk_size must be >= 1, got %llu",
                    (unsigned long long)cfg->chunk_size);
    if (cfg->window_mode != LDSC_WINDOW_CM && cfg->window_mode != LDSC_WINDOW_KB &&
        cfg->window_mode != LDSC_WINDOW_SNP)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG, "unknown window mode %d",
                    cfg->window_mode);
    if (!(cfg->window_value >= 0.0))
        return fail(err, err_len, LDSC_ERR_INVALID_ARG, "window value must be >= 0");
    if (cfg->precision != LDSC_PREC_INT8_FP64 && cfg->precision != LDSC_PREC_FP32)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG, "unknown precision %d",
                    cfg->precision);
    return LDSC_OK;
}

// The reference's inclusion rule (compute.rs:1164-1174, 1283-1299, 1855-1867,
// 2244-2257) reduced to one bound per SNP:
//   block_size[i] = i == bl[i] ? 0 : ceil((i - bl[i]) / c) * c
//   a_left(s)     = s - min(block_size[s], s)    (s a chunk start of the run)
// BB keeps every within-chunk pair, AB keeps window rows [a_left, s), so
// without masking pair (i<j) enters iff i >= a_left(chunk_start(j)).  Masking
// zeroes i < bl[j]; a_left(chunk_start(j)) <= bl[j], so masked lo[j] = bl[j].
void pair_bounds(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                 std::vector<uint64_t>& lo, bool* any_full) {
    std::vector<uint64_t> bl;
    bool full = block_lefts(snps, m, cfg->window_mode, cfg->window_value, bl);
    if (any_full) *any_full = full;
    lo.assign(m, 0);
    if (cfg->snp_level_masking) {
        lo = bl;
        return;
    }
    const uint64_t c = cfg->chunk_size;
    uint64_t run0 = 0;
    while (run0 < m) {
        uint64_t run1 = m;
        if (cfg->per_chromosome) {
            run1 = run0 + 1;
            while (run1 < m && snps[run1].chr == snps[run0].chr) ++run1;
        }
        for (uint64_t s = run0; s < run1; s += c) {
            const uint64_t sl = s - run0;  // chunk start, local to the run
            const uint64_t dist = s - bl[s];
            uint64_t bsz = dist == 0 ? 0 : ((dist + c - 1) / c) * c;
            if (bsz > sl) bsz = sl;
            const uint64_t a_left = s - bsz;
            const uint64_t e = s + c < run1 ? s + c : run1;
            for (uint64_t j = s; j < e; ++j) lo[j] = a_left;
        }
        run0 = run1;
    }
}

// fastrand 2.3.0 (wyrand, published algorithm): gen_u64 adds 0x2d358dccaa6c78a5
// to the state and folds the 128-bit product of s and s ^ 0x8bb84b93962eacc9;
// u32(..d) is Lemire's multiply-shift with rejection; bool() tests the low
// bit of gen_u32.  Call order of CountSketchProj::new: N buckets, then N signs.
static inline uint64_t wy_next(uint64_t& state) {
    const uint64_t s = state + 0x2d358dccaa6c78a5ULL;
    state = s;
    const unsigned __int128 t = (unsigned __int128)s * (unsigned __int128)(s ^ 0x8bb84b93962eacc9ULL);
    return (uint64_t)t ^ (uint64_t)(t >> 64);
}
static inline uint32_t wy_mod_u32(uint64_t& state, uint32_t n) {
    uint32_t r = (uint32_t)wy_next(state);
    uint64_t mm = (uint64_t)r * n;
    uint32_t lo = (uint32_t)mm;
    if (lo < n) {
        const uint32_t t = (uint32_t)(0u - n) % n;
        while (lo < t) {
            r = (uint32_t)wy_next(state);
            mm = (uint64_t)r * n;
            lo = (uint32_t)mm;
        }
    }
    return (uint32_t)(mm >> 32);
}
void countsketch_hash(uint64_t n, uint64_t d, uint64_t seed, uint32_t* bucket, int8_t* sign) {
    uint64_t st = seed;
    for (uint64_t i = 0; i < n; ++i) bucket[i] = wy_mod_u32(st, (uint32_t)d);
    for (uint64_t i = 0; i < n; ++i) sign[i] = ((uint8_t)wy_next(st) % 2 == 0) ? 1 : -1;
}

}  // namespace ldsc

extern "C" int ldsc_l2_block_left(const ldsc_snp* snps, uint64_t m, int32_t window_mode,
                                  double window_value, uint64_t* block_left_out,
                                  int32_t* any_full_chr_window, char* err, uint64_t err_len) {
    if ((!snps || !block_left_out) && m > 0)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL SNP or output array");
    if (window_mode != LDSC_WINDOW_CM && window_mode != LDSC_WINDOW_KB &&
        window_mode != LDSC_WINDOW_SNP)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "unknown window mode %d",
                          window_mode);
    std::vector<uint64_t> bl;
    bool full = ldsc::block_lefts(snps, m, window_mode, window_value, bl);
    if (m) std::memcpy(block_left_out, bl.data(), sizeof(uint64_t) * m);
    if (any_full_chr_window) *any_full_chr_window = full ? 1 : 0;
    return LDSC_OK;
}

extern "C" int ldsc_l2_pair_bounds(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                                   uint64_t* lo_out, char* err, uint64_t err_len) {
    int rc = ldsc::check_config(cfg, err, err_len);
    if (rc) return rc;
    if ((!snps || !lo_out) && m > 0)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL SNP or output array");
    std::vector<uint64_t> lo;
    ldsc::pair_bounds(snps, m, cfg, lo, nullptr);
    if (m) std::memcpy(lo_out, lo.data(), sizeof(uint64_t) * m);
    return LDSC_OK;
}

extern "C" int ldsc_l2_countsketch_hash(uint64_t n, uint64_t d, uint64_t seed,
                                        uint32_t* bucket, int8_t* sign, char* err,
                                        uint64_t err_len) {
    if (d == 0 || d > 0xFFFFFFFFull)
        return ldsc::fail(err, err_len, LDSC_ERR_SKETCH_RANGE, "sketch dimension %llu invalid",
                          (unsigned long long)d);
    if ((!bucket || !sign) && n > 0)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL output array");
    ldsc::countsketch_hash(n, d, seed, bucket, sign);
    return LDSC_OK;
}
