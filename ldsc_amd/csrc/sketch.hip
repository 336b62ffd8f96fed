// sketch.hip -- CountSketch path (placeholder until the kernels land).
#include "device.h"

namespace ldsc {
struct BandPlan;
int run_l2_sketch(const ldsc_panel*, const ldsc_snp*, uint64_t, const ldsc_l2_config*,
                  const std::vector<uint64_t>&, const BandPlan&, const double*, uint64_t,
                  double*, double*, ldsc_l2_perf*, Workspace*, char* err, uint64_t err_len) {
    return fail(err, err_len, LDSC_ERR_INVALID_ARG, "--sketch is not available yet");
}
}  // namespace ldsc

extern "C" int ldsc_l2_sketch_buckets(const ldsc_panel*, const uint64_t*, uint64_t, uint64_t,
                                      int32_t*, int32_t*, char* err, uint64_t err_len) {
    return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "sketch buckets not available yet");
}
