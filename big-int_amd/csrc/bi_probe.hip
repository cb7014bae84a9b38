// bi_probe.hip -- measurement helpers: the device's v_mad_u64_u32 issue rate
// (the roofline of the multiply, which uses no matrix cores) and the HBM
// bandwidth of a plain copy (the calibration kernel for HBM byte counters).
#include "../../include/bigint.h"
#include "bi_ctx.h"

#include <algorithm>

namespace bi {

// 16 independent mad+carry chains per lane in groups of four: the same
// instruction sequence as the multiply's inner loop (bi::mac4), with no
// memory traffic.
__global__ void __launch_bounds__(256) k_mac_rate(uint32_t seed, uint32_t iters, uint64_t* out) {
    uint32_t a[16];
    uint64_t acc[16];
    uint32_t cnt[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        a[j] = (threadIdx.x + j) * 2654435761u + seed;
        acc[j] = j;
        cnt[j] = 0;
    }
    uint32_t b = seed ^ 0x9e3779b9u;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < 16; j += 4) mac4(acc + j, cnt + j, b, a[j], a[j + 1], a[j + 2], a[j + 3]);
        b += 0x10001u;
    }
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) s ^= acc[j] + cnt[j];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst, uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace bi

using namespace bi;

extern "C" {

int bi_device_info(bi_ctx* ctx, int* cus, int* clock_khz, char* name, size_t name_len) {
    if (!ctx) return BI_ERR_ARG;
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if (cus) *cus = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    return BI_OK;
}

int bi_probe_mac_rate(bi_ctx* ctx, double* macs_per_s) {
    if (!ctx || !macs_per_s) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e);
    const uint32_t blocks = (uint32_t)p.multiProcessorCount * 8, threads = 256, iters = 8192;
    void* out;
    if ((rc = scratch(ctx, 7, (size_t)blocks * threads * 8, &out))) return rc;
    // the best of several launches after a warm-up, so that the clock has
    // settled at its loaded value
    for (int w = 0; w < 3; w++)
        hipLaunchKernelGGL(k_mac_rate, dim3(blocks), dim3(threads), 0, ctx->stream, 1u, iters, (uint64_t*)out);
    double best = 0;
    for (int rep = 0; rep < 6; rep++) {
        if ((e = hipEventRecord(ctx->ev0, ctx->stream)) != hipSuccess) return hip_fail(ctx, e);
        hipLaunchKernelGGL(k_mac_rate, dim3(blocks), dim3(threads), 0, ctx->stream, 2u + rep, iters, (uint64_t*)out);
        if ((e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess) return hip_fail(ctx, e);
        if ((e = hipEventSynchronize(ctx->ev1)) != hipSuccess) return hip_fail(ctx, e);
        float ms = 0;
        if ((e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1)) != hipSuccess) return hip_fail(ctx, e);
        best = std::max(best, (double)blocks * threads * iters * 16 / (ms * 1e-3));
    }
    *macs_per_s = best;
    return BI_OK;
}

int bi_copy_kernel(bi_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || !dst || !src || (bytes & 15) || ((((uintptr_t)dst) | ((uintptr_t)src)) & 15)) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    const uint64_t n4 = bytes / 16;
    uint64_t blocks = (n4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return BI_OK;
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, (const u32x4*)src, (u32x4*)dst, n4);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

}  // extern "C"
