// bi_api_helpers.hip -- extern "C" entry points of the helper kernels.
#include "../../include/bigint.h"
#include "bi_ctx.h"

using namespace bi;

static int helper_check(bi_ctx* ctx, uint32_t m, uint64_t n, const void* a, const void* b) {
    if (!ctx || m == 0) return BI_ERR_ARG;
    if (n && (!a || !b)) return BI_ERR_ARG;
    if ((uint64_t)m * n > (1ull << 40)) return BI_ERR_SHAPE;
    return set_device(ctx);
}

static int done(bi_ctx* ctx, hipError_t e) { return e == hipSuccess ? BI_OK : hip_fail(ctx, e); }

extern "C" {

int bi_sub_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs,
                   uint8_t* sign) {
    int rc = helper_check(ctx, m, n, as, bs);
    if (rc || n == 0) return rc;
    if (!rs || !sign) return BI_ERR_ARG;
    // the span kernel decides a < b per instance itself (a compare pass
    // first only for m below 17, where a span holds more than 63 instances)
    return done(ctx, add_uniform(ctx->sub_cfg_for(m), as, bs, rs, m, n, true, ctx->stream, sign));
}

int bi_cmp_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, int8_t* out) {
    int rc = helper_check(ctx, m, n, as, bs);
    if (rc || n == 0) return rc;
    if (!out) return BI_ERR_ARG;
    return done(ctx, topscan(0, as, bs, m, n, out, ctx->stream));
}

int bi_prec_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, uint32_t* out) {
    int rc = helper_check(ctx, m, n, as, as);
    if (rc || n == 0) return rc;
    if (!out) return BI_ERR_ARG;
    hipError_t e = topscan(2, as, nullptr, m, n, out, ctx->stream);
    if (e == hipSuccess) e = prec_ref(out, n, ctx->stream);
    return done(ctx, e);
}

int bi_prec_u64_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, uint32_t* out) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    int rc = helper_check(ctx, 2 * m, n, as, as);
    if (rc || n == 0) return rc;
    if (!out) return BI_ERR_ARG;
    hipError_t e = topscan(2, (const uint32_t*)as, nullptr, 2 * m, n, out, ctx->stream);
    if (e == hipSuccess) e = prec_ref(out, n, ctx->stream, true);
    return done(ctx, e);
}

// u64-limb views of the same bytes: m limbs = 2m u32 digits
int bi_sub_u64_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, const uint64_t* bs, uint64_t* rs,
                   uint8_t* sign) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return bi_sub_u32_dev(ctx, 2 * m, n, (const uint32_t*)as, (const uint32_t*)bs, (uint32_t*)rs, sign);
}

int bi_cmp_u64_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, const uint64_t* bs, int8_t* out) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return bi_cmp_u32_dev(ctx, 2 * m, n, (const uint32_t*)as, (const uint32_t*)bs, out);
}

int bi_shift_u64_dev(bi_ctx* ctx, uint32_t m, uint64_t n, int32_t s, const uint64_t* as, uint64_t* rs) {
    if (m > 0x7FFFFFFFu || s > 0x3FFFFFFF || s < -0x3FFFFFFF) return BI_ERR_SHAPE;
    return bi_shift_u32_dev(ctx, 2 * m, n, 2 * s, (const uint32_t*)as, (uint32_t*)rs);
}

int bi_shift_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, int32_t s, const uint32_t* as, uint32_t* rs) {
    int rc = helper_check(ctx, m, n, as, rs);
    if (rc || n == 0) return rc;
    const size_t bytes = (size_t)m * n * 4;
    const uint32_t* src = as;
    const char *a0 = (const char*)as, *r0 = (const char*)rs;
    if (a0 < r0 + bytes && r0 < a0 + bytes) {   // overlapping: stage the source
        void* tmp;
        if ((rc = scratch(ctx, 3, bytes, &tmp))) return rc;
        hipError_t e = hipMemcpyAsync(tmp, as, bytes, hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e);
        src = (const uint32_t*)tmp;
    }
    return done(ctx, shift(src, rs, m, n, s, nullptr, ctx->stream));
}

int bi_muld_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* ds, uint32_t* rs) {
    int rc = helper_check(ctx, m, n, as, ds);
    if (rc || n == 0) return rc;
    if (!rs) return BI_ERR_ARG;
    return done(ctx, muld(as, ds, rs, m, n, ctx->stream));
}

int bi_quod_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* ds, uint32_t* rs,
                    uint32_t* rem) {
    int rc = helper_check(ctx, m, n, as, ds);
    if (rc || n == 0) return rc;
    if (!rs) return BI_ERR_ARG;
    return done(ctx, quod(as, ds, rs, rem, m, n, ctx->stream));
}

}  // extern "C"
