// internal.h -- shared declarations of the ldsc_l2 library (not part of the ABI).
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ldsc_l2.h"

namespace ldsc {

// Writes a formatted message into the caller's error buffer and returns code.
inline int fail(char* err, uint64_t err_len, int code, const char* fmt, ...) {
    if (err && err_len > 0) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, (size_t)err_len, fmt, ap);
        va_end(ap);
    }
    return code;
}

// get_block_lefts_by_chr (src/l2/window.rs:32-90).  Returns any_full_chr_window.
bool block_lefts(const ldsc_snp* snps, uint64_t m, int mode, double value,
                 std::vector<uint64_t>& bl);

// Validates the window part of a config; returns LDSC_OK or an error code.
int check_config(const ldsc_l2_config* cfg, char* err, uint64_t err_len);

// Pair bound lo[j] (pair i<j included iff lo[j] <= i) for the call's SNP list.
// Chunk grid: global from 0, or restarted at each chromosome run when
// cfg->per_chromosome.  Also reports whether any chromosome run is covered by
// one window.
void pair_bounds(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                 std::vector<uint64_t>& lo, bool* any_full);

// fastrand 2.3.0 wyrand stream of CountSketchProj::new (compute.rs:179-192).
void countsketch_hash(uint64_t n, uint64_t d, uint64_t seed, uint32_t* bucket, int8_t* sign);

}  // namespace ldsc
