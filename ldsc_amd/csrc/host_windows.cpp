// host_windows.cpp -- host-side index work of the l2 path: LD windows, the
// per-SNP pair bound that encodes the reference's chunk/ring/mask inclusion
// rule, and the seeded CountSketch hash.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include <sched.h>

#include "internal.h"

namespace ldsc {

// ------------------------------------------------------------------------
// Host thread pool
// ------------------------------------------------------------------------
namespace {
class TaskPool {
  public:
    explicit TaskPool(int nt) {
        for (int i = 0; i < nt; ++i) std::thread([this] { worker(); }).detach();
        nthreads_ = nt;
    }
    int threads() const { return nthreads_; }
    // Runs f(0) .. f(n-1) on the caller and up to n-1 workers.  Only as many
    // workers as can help are woken, and the caller waits only for those
    // that joined: it closes the job once its own loop ends, and a worker
    // that wakes later stays out.
    void run(int n, const std::function<void(int)>& f) {
        std::unique_lock<std::mutex> busy(run_mu_, std::try_to_lock);
        if (!busy.owns_lock() || nthreads_ == 0 || n <= 1) {
            for (int i = 0; i < n; ++i) f(i);
            return;
        }
        const int want = std::min(n - 1, nthreads_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            ntasks_ = n;
            next_.store(0);
            slots_ = want;
            running_ = 0;
            ++gen_;
        }
        for (int i = 0; i < want; ++i) cv_.notify_one();
        for (int i; (i = next_.fetch_add(1)) < n;) f(i);
        std::unique_lock<std::mutex> lk(mu_);
        slots_ = 0;
        done_cv_.wait(lk, [this] { return running_ == 0; });
        job_ = nullptr;
    }

  private:
    void worker() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* f;
            int n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen && slots_ > 0; });
                seen = gen_;
                --slots_;
                ++running_;
                f = job_;
                n = ntasks_;
            }
            for (int i; (i = next_.fetch_add(1)) < n;) (*f)(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--running_ == 0) done_cv_.notify_one();
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)>* job_ = nullptr;
    int ntasks_ = 0, slots_ = 0, running_ = 0, nthreads_ = 0;
    std::atomic<int> next_{0};
    uint64_t gen_ = 0;
};

// CPUs this process may run on (its affinity mask, which a container or a
// batch slot narrows; hardware_concurrency counts every CPU of the node).
int usable_cpus() {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
    return (int)std::thread::hardware_concurrency();
}

// Pools are never destroyed: their workers outlive main.  The planning pool
// takes all usable CPUs but one, at most 15 workers.
TaskPool& pool() {
    static TaskPool* p = [] { return new TaskPool(std::min(std::max(usable_cpus() - 1, 0), 15)); }();
    return *p;
}

// Copies that run beside the host's planning (uploads) use a pool of their own.
TaskPool& copy_pool() {
    static TaskPool* p = [] {
        const unsigned hw = std::thread::hardware_concurrency();
        return new TaskPool((int)std::min<unsigned>(hw > 8 ? 3 : 1, 3));
    }();
    return *p;
}
}  // namespace

void parallel_tasks(int n, const std::function<void(int)>& f) { pool().run(n, f); }

int task_count(uint64_t items) {
    if (items < 32768) return 1;
    return (int)std::min<uint64_t>((uint64_t)pool().threads() + 1, items / 16384);
}

void parallel_tasks_side(int n, const std::function<void(int)>& f) { copy_pool().run(n, f); }

int copy_task_count(size_t bytes, bool side) {
    TaskPool& p = side ? copy_pool() : pool();
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)p.threads() + 1, bytes >> 18));
}

void parallel_copy(void* dst, const void* src, size_t bytes, bool side) {
    TaskPool& p = side ? copy_pool() : pool();
    const int T = copy_task_count(bytes, side);
    p.run(T, [&](int t) {
        // piece boundaries on 64-byte multiples
        const size_t a = (bytes * t / T) & ~(size_t)63;
        const size_t b = t + 1 == T ? bytes : (bytes * (t + 1) / T) & ~(size_t)63;
        std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
    });
}

// ------------------------------------------------------------------------
// LD windows
// ------------------------------------------------------------------------
static inline double coord_of(const ldsc_snp& s, bool cm) {
    return cm ? s.cm : (double)s.bp / 1000.0;
}

struct Run {
    uint64_t start, end;
};

// get_block_lefts_f64 (src/l2/window.rs:15-28) over i in [i0, i1) of the
// chromosome run [start, end) of the coordinates c: first j with
// |c_i - c_j| <= W, advancing while the distance is strictly greater.
static void block_lefts_scan(const double* c, uint64_t start, uint64_t i0, uint64_t i1,
                             uint64_t end, double W, uint64_t* out) {
    uint64_t j = start;
    for (uint64_t i = i0; i < i1; ++i) {
        const double ci = c[i];
        while (j < end && std::fabs(ci - c[j]) > W) ++j;
        out[i] = j;
    }
}

// First j in [start, i] with !(|c_i - c_j| > W), by binary search: valid
// where "distance > W" is a prefix property of j <= i.
static inline uint64_t first_in_window(const double* c, uint64_t start, uint64_t i, double W) {
    const double ci = c[i];
    uint64_t lo = start, hi = i;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (std::fabs(ci - c[mid]) > W) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One step of the scan from j for SNP i, testing four j at once: the advance
// is the length of the prefix of true tests.
static inline uint64_t advance_window(const double* c, double ci, uint64_t j, uint64_t end,
                                      double W) {
    for (;;) {
        if (j + 4 > end) {
            while (j < end && std::fabs(ci - c[j]) > W) ++j;
            return j;
        }
        const unsigned p0 = std::fabs(ci - c[j]) > W, p1 = std::fabs(ci - c[j + 1]) > W,
                       p2 = std::fabs(ci - c[j + 2]) > W, p3 = std::fabs(ci - c[j + 3]) > W;
        const unsigned q1 = p0 & p1, q2 = q1 & p2;
        const unsigned adv = p0 + q1 + q2 + (q2 & p3);
        j += adv;
        if (adv < 4) return j;
    }
}

// The same over a run whose coordinates never decrease (no NaN).  Then
// "distance > W" is a prefix property of j <= i (rounding is monotone) and
// false at j = i (W >= 0), so a scan may begin anywhere with a binary search
// for its first j.  Four such scans over consecutive quarters of [i0, i1)
// run interleaved, so that their dependent chains (each j waits for the one
// before) overlap in the core.  The result is the single scan's.
static void block_lefts_sorted(const double* c, uint64_t start, uint64_t i0, uint64_t i1,
                               uint64_t end, double W, uint64_t* out) {
    constexpr int S = 4;
    uint64_t is[S + 1], js[S];
    for (int q = 0; q <= S; ++q) is[q] = i0 + (i1 - i0) * q / S;
    for (int q = 0; q < S; ++q) js[q] = first_in_window(c, start, is[q], W);
    const uint64_t len = is[1] - is[0];  // the shortest quarter
    for (uint64_t t = 0; t < len; ++t) {
#pragma unroll
        for (int q = 0; q < S; ++q) {
            const uint64_t i = is[q] + t;
            js[q] = advance_window(c, c[i], js[q], end, W);
            out[i] = js[q];
        }
    }
    for (int q = 0; q < S; ++q)
        for (uint64_t i = is[q] + len; i < is[q + 1]; ++i) {
            js[q] = advance_window(c, c[i], js[q], end, W);
            out[i] = js[q];
        }
}

// Contiguous runs of one chromosome code (window.rs:41-49) of m >= 1 SNPs,
// found in parallel pieces: each task lists the run starts in its share.
// With `c`, the same pass stores each SNP's window coordinate there and sets
// `unsorted` for every run holding a coordinate that is not >= the one
// before it (a NaN counts).
static std::vector<Run> chromosome_runs(const ldsc_snp* snps, uint64_t m, bool cm = false,
                                        double* c = nullptr,
                                        std::vector<uint8_t>* unsorted = nullptr) {
    const int T = task_count(m);
    std::vector<std::vector<uint64_t>> starts(T), drops(T);
    parallel_tasks(T, [&](int t) {
        const uint64_t a = m * t / T, b = m * (t + 1) / T;
        const uint64_t k0 = std::max<uint64_t>(1, a);
        if (c) {
            if (a == 0) c[0] = coord_of(snps[0], cm);
            double prev = k0 - 1 >= a ? c[k0 - 1] : coord_of(snps[k0 - 1], cm);
            bool dropped = false;  // one drop per run piece is enough
            for (uint64_t k = k0; k < b; ++k) {
                const double ck = coord_of(snps[k], cm);
                c[k] = ck;
                if (snps[k].chr != snps[k - 1].chr) {
                    starts[t].push_back(k);
                    dropped = false;
                } else if (!(ck >= prev) && !dropped) {
                    drops[t].push_back(k);
                    dropped = true;
                }
                prev = ck;
            }
        } else {
            for (uint64_t k = k0; k < b; ++k)
                if (snps[k].chr != snps[k - 1].chr) starts[t].push_back(k);
        }
    });
    std::vector<Run> runs;
    uint64_t s0 = 0;
    for (const auto& v : starts)
        for (uint64_t k : v) {
            runs.push_back({s0, k});
            s0 = k;
        }
    runs.push_back({s0, m});
    if (unsorted) {
        unsorted->assign(runs.size(), 0);
        for (const auto& v : drops)
            for (uint64_t k : v) {
                // the run holding k: the last with start <= k
                const auto it = std::upper_bound(runs.begin(), runs.end(), k,
                                                 [](uint64_t x, const Run& r) { return x < r.start; });
                (*unsorted)[(size_t)(it - runs.begin()) - 1] = 1;
            }
    }
    return runs;
}

// `coord` (nullable): scratch for the window coordinates that the caller
// keeps with its capacity, so that a repeated call maps no fresh pages.
static bool block_lefts_runs(const ldsc_snp* snps, uint64_t m, int mode, double value,
                             std::vector<uint64_t>& bl, std::vector<Run>& runs,
                             std::vector<double>* coord) {
    bl.resize(m);
    runs.clear();
    if (m == 0) return false;
    bool any_full = false;
    if (mode == LDSC_WINDOW_SNP) {
        runs = chromosome_runs(snps, m);
        const uint64_t half = (uint64_t)value;
        for (const Run& r : runs)
            if (half >= r.end - r.start - 1) any_full = true;
        const int T = task_count(m);
        parallel_tasks(T, [&](int t) {
            const uint64_t a = m * t / T, b = m * (t + 1) / T;
            size_t r = 0;
            for (uint64_t i = a; i < b; ++i) {
                while (runs[r].end <= i) ++r;
                const uint64_t li = i - runs[r].start;
                bl[i] = runs[r].start + (li > half ? li - half : 0);
            }
        });
        return any_full;
    }
    // one pass for the coordinates, the runs and which runs are unsorted;
    // sorted runs are scanned in parallel pieces, unsorted ones each whole
    std::vector<double> local;
    std::vector<double>& c = coord ? *coord : local;
    if (c.size() < m) c.resize(m);
    std::vector<uint8_t> unsorted;
    runs = chromosome_runs(snps, m, mode == LDSC_WINDOW_CM, c.data(), &unsorted);
    const int T = task_count(m);
    parallel_tasks(T, [&](int t) {
        const uint64_t a = m * t / T, b = m * (t + 1) / T;
        for (size_t r = 0; r < runs.size(); ++r) {
            const Run& R = runs[r];
            if (unsorted[r] || R.end <= a || R.start >= b) continue;
            block_lefts_sorted(c.data(), R.start, std::max(a, R.start), std::min(b, R.end), R.end,
                               value, bl.data());
        }
    });
    std::vector<size_t> whole;
    for (size_t r = 0; r < runs.size(); ++r)
        if (unsorted[r]) whole.push_back(r);
    parallel_tasks((int)whole.size(), [&](int t) {
        const Run& R = runs[whole[t]];
        block_lefts_scan(c.data(), R.start, R.start, R.end, R.end, value, bl.data());
    });
    for (const Run& R : runs)
        if (bl[R.end - 1] == R.start) any_full = true;
    return any_full;
}

bool block_lefts(const ldsc_snp* snps, uint64_t m, int mode, double value,
                 std::vector<uint64_t>& bl, std::vector<double>* coord) {
    std::vector<Run> runs;
    return block_lefts_runs(snps, m, mode, value, bl, runs, coord);
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
    if (cfg->precision != LDSC_PREC_EXACT && cfg->precision != LDSC_PREC_FP32 &&
        cfg->precision != LDSC_PREC_EXACT_I8)
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
                 std::vector<uint64_t>& lo, bool* any_full, std::vector<double>* coord) {
    // lo starts as block_left and is rewritten in place chunk by chunk: each
    // chunk reads block_left only at its own start, before writing over it,
    // so chunks are independent tasks (owned by the task holding their start)
    std::vector<Run> runs;
    const bool full =
        block_lefts_runs(snps, m, cfg->window_mode, cfg->window_value, lo, runs, coord);
    if (any_full) *any_full = full;
    if (cfg->snp_level_masking || m == 0) return;
    const uint64_t c = cfg->chunk_size;
    // chunk grids: one per chromosome run, or one over all SNPs
    if (!cfg->per_chromosome) runs.assign(1, Run{0, m});
    const int T = task_count(m);
    parallel_tasks(T, [&](int t) {
        const uint64_t a = m * t / T, b = m * (t + 1) / T;
        for (const Run& rr : runs) {
            if (rr.end <= a || rr.start >= b) continue;
            // first chunk start of the run at or after a
            uint64_t s = rr.start;
            if (a > s) s += (a - s + c - 1) / c * c;
            for (; s < rr.end && s < b; s += c) {
                const uint64_t sl = s - rr.start;  // chunk start, local to the run
                const uint64_t dist = s - lo[s];
                uint64_t bsz = dist == 0 ? 0 : ((dist + c - 1) / c) * c;
                if (bsz > sl) bsz = sl;
                const uint64_t a_left = s - bsz;
                const uint64_t e = s + c < rr.end ? s + c : rr.end;
                for (uint64_t j = s; j < e; ++j) lo[j] = a_left;
            }
        }
    });
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
