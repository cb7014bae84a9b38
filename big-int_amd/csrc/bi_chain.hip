// bi_chain.hip -- fused chains of adds and of products, kept on chip between
// the operations (one HBM read of the operands, one write of the result).
//
// Reference: the in-kernel benchmark chains of the thesis,
//   baddKer1Bench / baddKer2Bench / baddKer3Bench (cuda/ker-add.cu.h:88-112,
//   175-203, 268-295): r1 = a + b, then r = r1 + r, p additions in all, i.e.
//   r = p (a + b) mod B^m;
//   convMul1Bench .. convMul5Bench (cuda/ker-mul.cu.h:223-250 and twins):
//   r1 = a b, then r = r1 r, p products in all, i.e. r = (a b)^p mod B^m.
// The reference's v2/v3 add chains alias their operand and result registers
// and lose carries (the values they compute are not p (a + b)); here every
// chain computes the closed form above.
#include <algorithm>

#include "../../include/bigint.h"
#include "bi_addcore.h"
#include "bi_ctx.h"
#include "bi_kernels.h"

namespace bi {

// One wave per span of ipw whole instances of m digits held in NT tiles of
// 256 digits (4 per lane) in registers for the whole chain.  SEG: the span
// holds several instances, whose starts cut the carry chain; with one
// instance (ipw == 1) every pass starts at it with carry 0, so the segment
// flags are all zero and their handling compiles away.
template <bool VEC, int NT, bool SEG>
__global__ void __launch_bounds__(256) k_add_chain(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                   uint32_t* __restrict__ rs, uint64_t n, uint32_t m, uint32_t ipw,
                                                   uint32_t p, uint64_t nspans) {
    const uint32_t lane = lane_id();
    const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (w >= nspans) return;
    const uint64_t i0 = w * ipw;
    const uint64_t i1 = std::min<uint64_t>(i0 + ipw, n);
    const uint64_t s0 = i0 * m, s1 = i1 * m;
    u32x4 r1[NT], r[NT];
    uint32_t fl[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const uint64_t f = s0 + 256 * t + 4 * lane;
        r1[t] = load_grp<VEC, true>(as, f, s1);
        r[t] = load_grp<VEC, true>(bs, f, s1);
        fl[t] = (SEG && f < s1) ? uniform_flags(f, s0, m) : 0u;
    }
    // r1 = a + b
    uint32_t carry = 0;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        u32x4 s;
        bool g, pr;
        lane_pre<false>(r1[t], r[t], fl[t], s, g, pr);
        uint32_t cout;
        const uint32_t e = wave_carry(g, pr, carry, cout);
        carry = cout;
        r1[t] = lane_post<false>(s, fl[t], e);
        r[t] = r1[t];
    }
    // r = r1 + r, p - 1 more times
    for (uint32_t i = 1; i < p; i++) {
        carry = 0;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            u32x4 s;
            bool g, pr;
            lane_pre<false>(r1[t], r[t], fl[t], s, g, pr);
            uint32_t cout;
            const uint32_t e = wave_carry(g, pr, carry, cout);
            carry = cout;
            r[t] = lane_post<false>(s, fl[t], e);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const uint64_t f = s0 + 256 * t + 4 * lane;
        store_grp<VEC, true>(rs, f, r[t], s0, s1);
    }
}

hipError_t add_chain_uniform(const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m, uint64_t n, uint32_t p,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (m == 0 || m > kAddChainMaxDigits || p == 0) return hipErrorInvalidValue;
    uint32_t NT = 1;
    while (NT * 256 < m) NT *= 2;
    const uint32_t ipw = std::max<uint32_t>(1, NT * 256 / m);
    const uint64_t nspans = (n + ipw - 1) / ipw;
    const bool vec = (m % 4 == 0) && ((((uintptr_t)as) | ((uintptr_t)bs) | ((uintptr_t)rs)) & 15u) == 0;
    const dim3 grid((uint32_t)((nspans + 3) / 4)), block(256);
#define BI_CHAIN_LAUNCH(K)                                                                                   \
    do {                                                                                                     \
        if (ipw > 1) {                                                                                       \
            if (vec) hipLaunchKernelGGL((k_add_chain<true, K, true>), grid, block, 0, st, as, bs, rs, n, m, ipw, p, nspans); \
            else hipLaunchKernelGGL((k_add_chain<false, K, true>), grid, block, 0, st, as, bs, rs, n, m, ipw, p, nspans); \
        } else {                                                                                             \
            if (vec) hipLaunchKernelGGL((k_add_chain<true, K, false>), grid, block, 0, st, as, bs, rs, n, m, ipw, p, nspans); \
            else hipLaunchKernelGGL((k_add_chain<false, K, false>), grid, block, 0, st, as, bs, rs, n, m, ipw, p, nspans); \
        }                                                                                                    \
    } while (0)
    switch (NT) {
        case 1: BI_CHAIN_LAUNCH(1); break;
        case 2: BI_CHAIN_LAUNCH(2); break;
        case 4: BI_CHAIN_LAUNCH(4); break;
        default: BI_CHAIN_LAUNCH(8); break;
    }
#undef BI_CHAIN_LAUNCH
    return hipGetLastError();
}

}  // namespace bi

using namespace bi;

extern "C" {

// rs = p (as + bs) mod B^m per instance (the reference's "ten additions" for
// p = 10), computed as p chained additions.  m <= 2048: one kernel, the
// chain in registers; larger m: one add launch per link (r1 kept in scratch).
int bi_add_chain_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, uint32_t p, const uint32_t* as, const uint32_t* bs,
                         uint32_t* rs) {
    if (!ctx || m == 0 || p == 0) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    if (!as || !bs || !rs) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipError_t e;
    if (m <= kAddChainMaxDigits) {
        e = add_chain_uniform(as, bs, rs, m, n, p, ctx->stream);
        return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
    }
    e = add_uniform(ctx->add_cfg_for(m), as, bs, rs, m, n, false, ctx->stream, nullptr);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if (p == 1) return BI_OK;
    void* t1;
    if ((rc = scratch(ctx, 6, (size_t)n * m * 4, &t1))) return rc;
    e = hipMemcpyAsync(t1, rs, (size_t)n * m * 4, hipMemcpyDeviceToDevice, ctx->stream);
    for (uint32_t i = 1; i < p && e == hipSuccess; i++)
        e = add_uniform(ctx->add_cfg_for(m), (const uint32_t*)t1, rs, rs, m, n, false, ctx->stream, nullptr);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

// rs = (as bs)^p mod B^m per instance (the reference's "six multiplications"
// for p = 6): r1 = a b, r_i = r1 r_{i-1}.  m <= 512: one kernel, the chain in
// LDS; larger m (<= 4096): one multiply launch per link.
int bi_mul_chain_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, uint32_t p, const uint32_t* as, const uint32_t* bs,
                         uint32_t* rs) {
    if (!ctx || m == 0 || p == 0) return BI_ERR_ARG;
    if (m > kMulMaxDigits) return BI_ERR_SHAPE;
    if (n == 0) return BI_OK;
    if (!as || !bs || !rs) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    uint32_t* snaps;
    uint64_t snap_words;
    if ((rc = mul_snaps(ctx, &snaps, &snap_words))) return rc;
    hipError_t e;
    if (m <= kMulChainMaxDigits) {
        e = mul_chain_uniform(ctx->mul_cfg, as, bs, rs, m, n, p, ctx->stream, snaps, snap_words,
                              (uint32_t)ctx->num_cus);
        return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
    }
    e = mul_uniform(ctx->mul_cfg, as, bs, rs, m, n, ctx->stream, snaps, snap_words, (uint32_t)ctx->num_cus);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if (p == 1) return BI_OK;
    void* t1;
    if ((rc = scratch(ctx, 6, (size_t)n * m * 4, &t1))) return rc;
    e = hipMemcpyAsync(t1, rs, (size_t)n * m * 4, hipMemcpyDeviceToDevice, ctx->stream);
    for (uint32_t i = 1; i < p && e == hipSuccess; i++)
        e = mul_uniform(ctx->mul_cfg, (const uint32_t*)t1, rs, rs, m, n, ctx->stream, snaps, snap_words,
                        (uint32_t)ctx->num_cus);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

}  // extern "C"
