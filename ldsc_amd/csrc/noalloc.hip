// noalloc.hip -- ldsc_l2_compute without device or pinned-host allocation.
//
// For callers that cannot allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Every device buffer of
// the call (tiled SNP slice, stats, pair bounds, band plan, tile partials,
// A_eff, results) is carved from one file-scope device array, `scratch`, of
// kNoAllocBytes; host staging is pageable memory kept from call to call.
// The kernels and host phases are those of ldsc_l2_compute (run_l2 on a
// workspace in arena mode), so results are bit-identical to it.
//
// Set-up, once per device on the first call there: the device address of
// `scratch` (hipGetSymbolAddress; host code never passes the array's own
// name to a kernel), and the workspace's streams and events.  Calls on one
// device are serialized by a lock: they share that device's array, whose
// contents are the last call's (nothing in it needs to start at zero: every
// buffer is written before it is read, as in the allocating path).
#include <mutex>

#include "band_common.h"

namespace ldsc {

constexpr size_t kNoAllocBytes = 256ull << 20;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocBytes];

namespace {
struct NoAllocDevice {
    std::once_flag once;
    hipError_t setup = hipSuccess;
    std::mutex mu;
    Arena arena;
    Workspace ws;
};
NoAllocDevice g_noalloc[kNoAllocDevices];

void noalloc_setup(NoAllocDevice& d, int dev) {
    void* ptr = nullptr;
    hipError_t e = hipSetDevice(dev);
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && !ptr) e = hipErrorInvalidSymbol;
    if (e == hipSuccess) e = d.ws.init(dev);
    if (e == hipSuccess) {
        d.arena.base = static_cast<char*>(ptr);
        d.arena.cap = kNoAllocBytes;
        d.arena.off = 0;
        d.ws.use_arena(&d.arena);
    }
    d.setup = e;
}
}  // namespace

}  // namespace ldsc

using namespace ldsc;

extern "C" int ldsc_l2_compute_noalloc(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m,
                                       const ldsc_l2_config* cfg, const double* annot, uint64_t K,
                                       double* l2_out, double* maf_out, ldsc_l2_perf* perf,
                                       ldsc_progress_cb cb, void* cb_ctx, char* err,
                                       uint64_t err_len) {
    if (!panel) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "panel is NULL");
    int rc = check_config(cfg, err, err_len);
    if (rc) return rc;
    if (cfg->sketch_d)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG,
                    "ldsc_l2_compute_noalloc: the CountSketch path (--sketch) allocates its "
                    "per-panel plan; use ldsc_l2_compute");
    if (m > 0 && !snps) return fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL SNP array");
    if (K < 1) return fail(err, err_len, LDSC_ERR_ANNOT, "annotation column count must be >= 1");
    if (panel->device < 0 || panel->device >= kNoAllocDevices)
        return fail(err, err_len, LDSC_ERR_INVALID_ARG, "device %d out of range", panel->device);
    // the call's own device bytes, from its band, against the fixed maximum
    if (m > 0 && m < (1ull << 31)) {
        std::vector<uint64_t> lo;
        const uint64_t tiles = band_tile_count(snps, m, cfg, lo);
        const size_t need = l2_workspace_bytes(panel, m, cfg, K, tiles, annot != nullptr);
        if (need > kNoAllocBytes)
            return fail(err, err_len, LDSC_ERR_OOM,
                        "ldsc_l2_compute_noalloc: this call needs ~%llu device bytes, over the "
                        "fixed %llu-byte scratch (use ldsc_l2_compute)",
                        (unsigned long long)need, (unsigned long long)kNoAllocBytes);
    }
    NoAllocDevice& d = g_noalloc[panel->device];
    std::call_once(d.once, [&] { noalloc_setup(d, panel->device); });
    if (d.setup != hipSuccess)
        return fail(err, err_len, LDSC_ERR_DEVICE, "ldsc_l2_compute_noalloc set-up failed: %s",
                    hipGetErrorString(d.setup));
    std::lock_guard<std::mutex> lk(d.mu);
    // A call that failed part-way can leave copies in flight on the side
    // streams; the array is carved again from its start, in another layout
    // when the shape differs, so they must end first.
    for (hipStream_t s : {d.ws.stream, d.ws.copy_stream, d.ws.up_stream, d.ws.plan_stream}) {
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess)
            return fail(err, err_len, LDSC_ERR_DEVICE, "ldsc_l2_compute_noalloc: %s",
                        hipGetErrorString(e));
    }
    d.ws.reset_arena();
    return run_l2(panel, snps, m, cfg, annot, K, l2_out, maf_out, perf, cb, cb_ctx, &d.ws, err,
                  err_len);
}
