// calib.hip -- measured peaks of the MI355X for the roofline figures of
// bench.py: the sustained rate of the MFMA forms the band kernels use (e2m1,
// int8, FP32) with every SIMD busy, and the device-to-device copy bandwidth
// of HBM.  Not part of the LD-score ABI: these kernels only time the
// hardware (synthetic register operands, one result store per lane).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FORM 0: v_mfma_scale_f32_32x32x64_f8f6f4 on e2m1 operands (131,072 flops)
// FORM 1: v_mfma_i32_32x32x32_i8 (65,536 ops)
// FORM 2: v_mfma_f32_32x32x2f32 (4,096 flops)
// 16 independent 32x32 accumulators per wave, one wave per SIMD.
template <int FORM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_mfma_rate(int iters, uint32_t seed, float* out, unsigned long long* cycles) {
    const int lane = threadIdx.x;
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = (seed * (uint32_t)(lane + 1) * (uint32_t)(i + 3)) & 0x33333333u;
    const unsigned long long t0 = __builtin_readcyclecounter();
    float s = 0.0f;
    if constexpr (FORM == 0) {
        v16f acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][r] = 0.0f;
        const v4i xa = {(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
        const v4i xb = {(int)w[4], (int)w[5], (int)w[6], (int)w[7]};
        const v8i A = __builtin_shufflevector(xa, xa, 0, 1, 2, 3, -1, -1, -1, -1);
        const v8i B = __builtin_shufflevector(xb, xb, 0, 1, 2, 3, -1, -1, -1, -1);
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                acc[i] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    A, B, acc[i], 4, 4, 0, 0x7f7f7f7f, 0, 0x7f7f7f7f);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) s += acc[i][r];
    } else if constexpr (FORM == 1) {
        v16i acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][r] = 0;
        const v4i A = {(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
        const v4i B = {(int)w[4], (int)w[5], (int)w[6], (int)w[7]};
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                acc[i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B, acc[i], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) s += (float)acc[i][r];
    } else {
        v16f acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][r] = 0.0f;
        const float a = (float)(w[0] & 0xFF) * 0.01f, b = (float)(w[1] & 0xFF) * 0.01f;
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                acc[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[i], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) s += acc[i][r];
    }
    const unsigned long long t1 = __builtin_readcyclecounter();
    out[(size_t)blockIdx.x * 64 + lane] = s;
    if (lane == 0) cycles[blockIdx.x] = t1 - t0;
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Streaming copy, grid-stride: four 16-byte loads per lane in flight before
// their stores, non-temporal (the data is touched once).
__global__ __launch_bounds__(256) void k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                              size_t n) {
    const size_t stride = (size_t)gridDim.x * 256;
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

constexpr double kOps[3] = {32.0 * 32 * 64 * 2, 32.0 * 32 * 32 * 2, 32.0 * 32 * 2 * 2};

}  // namespace

extern "C" {

// Sustained MFMA rate of `form` (0 e2m1, 1 int8, 2 FP32) on `device`, with
// one wave per SIMD on every CU running `iters` x 16 MFMAs.  Writes the rate
// in TFLOP/s (TOP/s for int8) and the mean shader clock in GHz implied by
// the waves' cycle counts.  Returns 0 or a hipError_t code.
int ldsc_calib_mfma(int device, int form, int iters, double* tflops, double* ghz) {
    if (form < 0 || form > 2 || iters < 1) return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return (int)e;
    const int blocks = prop.multiProcessorCount * 4;
    float* out = nullptr;
    unsigned long long* cyc = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](hipError_t r) {
        if (out) (void)hipFree(out);
        if (cyc) (void)hipFree(cyc);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return (int)r;
    };
    if ((e = hipMalloc(&out, (size_t)blocks * 64 * sizeof(float))) != hipSuccess) return done(e);
    if ((e = hipMalloc(&cyc, (size_t)blocks * sizeof(unsigned long long))) != hipSuccess) return done(e);
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        return done(e);
    auto kfn = form == 0 ? k_mfma_rate<0> : form == 1 ? k_mfma_rate<1> : k_mfma_rate<2>;
    // warm up the clocks with one full-length run, then time a second one
    hipLaunchKernelGGL(kfn, dim3(blocks), dim3(64), 0, 0, iters, 7u, out, cyc);
    if ((e = hipEventRecord(e0, 0)) != hipSuccess) return done(e);
    hipLaunchKernelGGL(kfn, dim3(blocks), dim3(64), 0, 0, iters, 7u, out, cyc);
    if ((e = hipGetLastError()) != hipSuccess) return done(e);
    if ((e = hipEventRecord(e1, 0)) != hipSuccess) return done(e);
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return done(e);
    float ms = 0.0f;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return done(e);
    std::vector<unsigned long long> c(blocks);
    if ((e = hipMemcpy(c.data(), cyc, c.size() * sizeof(c[0]), hipMemcpyDeviceToHost)) != hipSuccess)
        return done(e);
    double cs = 0.0;
    for (auto x : c) cs += (double)x;
    cs /= (double)blocks;
    if (tflops) *tflops = (double)blocks * iters * 16 * kOps[form] / (ms * 1e-3) / 1e12;
    if (ghz) *ghz = cs / (ms * 1e-3) / 1e9;
    return done(hipSuccess);
}

// Device-to-device copy bandwidth over `bytes` (read + written bytes per
// second, GB/s): the best of `reps` timed copies (after one warm-up) by the
// streaming kernel above and by the runtime's hipMemcpyAsync.
int ldsc_calib_copy(int device, uint64_t bytes, int reps, double* gbs) {
    if (bytes < 16 || reps < 1) return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return (int)e;
    const size_t n = (size_t)(bytes / 16);
    u32x4 *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](hipError_t r) {
        if (src) (void)hipFree(src);
        if (dst) (void)hipFree(dst);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return (int)r;
    };
    if ((e = hipMalloc(&src, n * 16)) != hipSuccess) return done(e);
    if ((e = hipMalloc(&dst, n * 16)) != hipSuccess) return done(e);
    if ((e = hipMemset(src, 1, n * 16)) != hipSuccess) return done(e);
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        return done(e);
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return done(e);
    const unsigned grid = (unsigned)prop.multiProcessorCount * 8;
    double best = 0.0;
    for (int r = 0; r <= reps; ++r) {
        if ((e = hipEventRecord(e0, 0)) != hipSuccess) return done(e);
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, src, dst, n);
        if ((e = hipGetLastError()) != hipSuccess) return done(e);
        if ((e = hipEventRecord(e1, 0)) != hipSuccess) return done(e);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) return done(e);
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return done(e);
        if (r > 0) best = std::max(best, 2.0 * (double)n * 16 / (ms * 1e-3) / 1e9);
    }
    // the runtime's own device-to-device copy over the same buffers
    for (int r = 0; r <= reps; ++r) {
        if ((e = hipEventRecord(e0, 0)) != hipSuccess) return done(e);
        if ((e = hipMemcpyAsync(dst, src, n * 16, hipMemcpyDeviceToDevice, 0)) != hipSuccess)
            return done(e);
        if ((e = hipEventRecord(e1, 0)) != hipSuccess) return done(e);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) return done(e);
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return done(e);
        if (r > 0) best = std::max(best, 2.0 * (double)n * 16 / (ms * 1e-3) / 1e9);
    }
    if (gbs) *gbs = best;
    return done(hipSuccess);
}

}  // extern "C"
