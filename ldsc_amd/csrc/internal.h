// internal.h -- shared declarations of the ldsc_l2 library (not part of the ABI).
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <functional>
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

// Runs f(0) .. f(n-1) on a small pool of persistent host threads and the
// calling thread; returns when all are done.  Calls that find the pool busy
// (another thread's call in progress) run their tasks inline.
void parallel_tasks(int n, const std::function<void(int)>& f);

// Number of tasks worth splitting `items` independent items into.
int task_count(uint64_t items);

// memcpy of `bytes` split over the task pool (pieces of at least 256 KiB);
// `side` uses a separate small pool, for copies that run while the calling
// thread's own work keeps the main pool busy.
void parallel_copy(void* dst, const void* src, size_t bytes, bool side = false);
// n tasks on the side pool (the one parallel_copy(.., side = true) uses).
void parallel_tasks_side(int n, const std::function<void(int)>& f);
// Tasks worth running for a copy of `bytes` on the main or the side pool.
int copy_task_count(size_t bytes, bool side);

// get_block_lefts_by_chr (src/l2/window.rs:32-90).  Returns any_full_chr_window.
// `coord` (nullable) is scratch for the window coordinates, kept by the
// caller with its capacity from call to call.
bool block_lefts(const ldsc_snp* snps, uint64_t m, int mode, double value,
                 std::vector<uint64_t>& bl, std::vector<double>* coord = nullptr);

// Validates the window part of a config; returns LDSC_OK or an error code.
int check_config(const ldsc_l2_config* cfg, char* err, uint64_t err_len);

// Pair bound lo[j] (pair i<j included iff lo[j] <= i) for the call's SNP list.
// Chunk grid: global from 0, or restarted at each chromosome run when
// cfg->per_chromosome.  Also reports whether any chromosome run is covered by
// one window.
void pair_bounds(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                 std::vector<uint64_t>& lo, bool* any_full,
                 std::vector<double>* coord = nullptr);

// Gather map in lane-major layout: map_t[e * n16 + q] is the FAM row of
// output member 16q + e (-1 past the last member), so that the 16 loads a
// thread makes for its output dword are each coalesced across the wave.
inline std::vector<int32_t> lane_major_map(const std::vector<int32_t>& rows) {
    const uint64_t n = rows.size(), n16 = (n + 15) / 16;
    std::vector<int32_t> t(16 * n16, -1);
    for (uint64_t k = 0; k < n; ++k) t[(k % 16) * n16 + k / 16] = rows[k];
    return t;
}

// fastrand 2.3.0 wyrand stream of CountSketchProj::new (compute.rs:179-192).
void countsketch_hash(uint64_t n, uint64_t d, uint64_t seed, uint32_t* bucket, int8_t* sign);

}  // namespace ldsc
