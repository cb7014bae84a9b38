// bi_gmpio.hip -- GMP-compatible export / import of batches: the format of
// mpz_export / mpz_import with order -1 (least significant word first),
// native (little) endian, no nails, words of 4 or 8 bytes -- the format the
// reference uses to hand integers to GMP and back (cuda/helpers.h:113,
// 167-177: mpz_import / mpz_export with order -1 and 4-byte words).
//
// Export of a uniform [n][m] batch is two calls: the word counts (what
// mpz_export would report: the significant words, 0 for zero) and their
// exclusive prefix (device offsets, n+1 entries), then the packed words, each
// integer's significant words back to back (an odd digit count in 8-byte
// words gets a zero high half, as mpz_export writes it).  Import reads such a
// packed stream into a uniform batch, zero-padded, truncated mod B^m.
#include "../../include/bigint.h"
#include "bi_ctx.h"
#include "bi_kernels.h"

namespace bi {

__global__ void __launch_bounds__(256) k_export_counts(const uint32_t* __restrict__ sig, uint32_t wsize, uint64_t n,
                                                       uint64_t* __restrict__ offs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) offs[0] = 0;
    if (i < n) {
        const uint32_t s = sig[i];
        offs[i + 1] = wsize == 8 ? (s + 1) / 2 : s;
    }
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t u = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += u;
    }
    return v;
}

// In-place inclusive scan of blocks of 1024 values; block totals to sums.
__global__ void __launch_bounds__(1024) k_scan_blocks(uint64_t* __restrict__ x, uint64_t n, uint64_t* __restrict__ sums) {
    __shared__ uint64_t wt[16];
    const uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    const uint32_t wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    uint64_t v = i < n ? x[i] : 0;
    v = wave_incl_scan(v);
    if (lane == 63) wt[wv] = v;
    __syncthreads();
    if (wv == 0) {
        uint64_t t = lane < 16 ? wt[lane] : 0;
        t = wave_incl_scan(t);
        if (lane < 16) wt[lane] = t;
    }
    __syncthreads();
    if (wv) v += wt[wv - 1];
    if (i < n) x[i] = v;
    if (threadIdx.x == 1023 && sums) sums[blockIdx.x] = v;
}

// Inclusive scan of the block totals by one workgroup, 1024 at a time.
__global__ void __launch_bounds__(1024) k_scan_sums(uint64_t* __restrict__ sums, uint64_t nb) {
    __shared__ uint64_t wt[16];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const uint32_t wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (uint64_t base = 0; base < nb; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        uint64_t v = i < nb ? sums[i] : 0;
        v = wave_incl_scan(v);
        if (lane == 63) wt[wv] = v;
        __syncthreads();
        if (wv == 0) {
            uint64_t t = lane < 16 ? wt[lane] : 0;
            t = wave_incl_scan(t);
            if (lane < 16) wt[lane] = t;
        }
        __syncthreads();
        if (wv) v += wt[wv - 1];
        v += carry;
        if (i < nb) sums[i] = v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = v;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) k_scan_add(uint64_t* __restrict__ x, uint64_t n, const uint64_t* __restrict__ sums) {
    const uint64_t i = (uint64_t)blockIdx.x * 1024 + threadIdx.x;
    if (blockIdx.x > 0 && i < n) x[i] += sums[blockIdx.x - 1];
}

// one wave per instance
// One wave per instance, 256 digits a wave per step: the packed side starts
// anywhere, so both sides move as single words, lane-consecutive (coalesced),
// four words a lane in flight, nontemporal on the [n][m] side.
__global__ void __launch_bounds__(256) k_export(const uint32_t* __restrict__ xs, uint32_t m, uint64_t n, uint32_t wsize,
                                                const uint64_t* __restrict__ offs, uint32_t* __restrict__ out) {
    const uint32_t lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (i >= n) return;
    const uint32_t q = wsize / 4;
    const uint64_t c32 = (offs[i + 1] - offs[i]) * q;
    const uint32_t* x = xs + i * m;
    uint32_t* o = out + offs[i] * q;
    for (uint64_t k0 = lane; k0 < c32; k0 += 4 * kWave) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t k = k0 + (uint64_t)j * kWave;
            v[j] = k < m ? __builtin_nontemporal_load(x + k) : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t k = k0 + (uint64_t)j * kWave;
            if (k < c32) o[k] = v[j];
        }
    }
}

__global__ void __launch_bounds__(256) k_import(const uint32_t* __restrict__ in, uint32_t m, uint64_t n, uint32_t wsize,
                                                const uint64_t* __restrict__ offs, uint32_t* __restrict__ xs) {
    const uint32_t lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (i >= n) return;
    const uint32_t q = wsize / 4;
    const uint64_t c32 = (offs[i + 1] - offs[i]) * q;
    const uint32_t* src = in + offs[i] * q;
    uint32_t* x = xs + i * m;
    for (uint32_t k0 = lane; k0 < m; k0 += 4 * kWave) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = k0 + j * kWave;
            v[j] = k < c32 ? src[k] : 0u;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = k0 + j * kWave;
            if (k < m) __builtin_nontemporal_store(v[j], x + k);
        }
    }
}

// in-place inclusive scan of x[0..n)
static hipError_t scan_u64(bi_ctx* ctx, uint64_t* x, uint64_t n) {
    if (n == 0) return hipSuccess;
    const uint64_t nb = (n + 1023) / 1024;
    void* sums = nullptr;
    if (nb > 1 && scratch(ctx, 9, nb * 8, &sums)) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_scan_blocks, dim3((uint32_t)nb), dim3(1024), 0, ctx->stream, x, n, (uint64_t*)sums);
    if (nb > 1) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, ctx->stream, (uint64_t*)sums, nb);
        hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)nb), dim3(1024), 0, ctx->stream, x, n, (const uint64_t*)sums);
    }
    return hipGetLastError();
}

}  // namespace bi

using namespace bi;

extern "C" {

int bi_export_sizes_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* xs, uint32_t wsize, uint64_t* offs) {
    if (!ctx || m == 0 || !offs || (n && !xs) || (wsize != 4 && wsize != 8)) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    void* sig;
    if ((rc = scratch(ctx, 8, (size_t)(n ? n : 1) * 4, &sig))) return rc;
    hipError_t e = topscan(2, xs, nullptr, m, n, sig, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_export_counts, dim3((uint32_t)((n + 256) / 256)), dim3(256), 0, ctx->stream,
                           (const uint32_t*)sig, wsize, n, offs);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = scan_u64(ctx, offs + 1, n);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_export_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* xs, uint32_t wsize, const uint64_t* offs,
                      void* out) {
    if (!ctx || m == 0 || (n && (!xs || !offs || !out)) || (wsize != 4 && wsize != 8)) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_export, dim3((uint32_t)((n + 3) / 4)), dim3(256), 0, ctx->stream, xs, m, n, wsize, offs,
                       (uint32_t*)out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_import_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const void* in, uint32_t wsize, const uint64_t* offs,
                      uint32_t* xs) {
    if (!ctx || m == 0 || (n && (!xs || !offs || !in)) || (wsize != 4 && wsize != 8)) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_import, dim3((uint32_t)((n + 3) / 4)), dim3(256), 0, ctx->stream, (const uint32_t*)in, m, n,
                       wsize, offs, xs);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

}  // extern "C"
