// bi_api.hip -- the extern "C" boundary of libbigint (declared in
// include/bigint.h): context, plumbing and the plain batched entry points.
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

#include "../../include/bigint.h"
#include "bi_ctx.h"

using namespace bi;

namespace bi {

int hip_fail(bi_ctx* ctx, hipError_t e) {
    if (ctx) {
        strncpy(ctx->last_err, hipGetErrorString(e), sizeof(ctx->last_err) - 1);
        ctx->last_err[sizeof(ctx->last_err) - 1] = 0;
    }
    (void)hipGetLastError();
    return BI_ERR_HIP;
}

int set_device(bi_ctx* ctx) {
    hipError_t e = hipSetDevice(ctx->device);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

// Grow-only device scratch: slot i is reused across calls on this context.
int scratch(bi_ctx* ctx, int slot, size_t bytes, void** out) {
    if (slot < 0 || slot >= kScratchSlots) return BI_ERR_ARG;
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch[slot]) {
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return hip_fail(ctx, e);
            (void)hipFree(ctx->scratch[slot]);
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return BI_ERR_NOMEM;
        }
        ctx->scratch[slot] = p;
        ctx->scratch_bytes[slot] = bytes;
    }
    *out = ctx->scratch[slot];
    return BI_OK;
}

uint32_t resident_per_cu(const void* fn, uint32_t block, size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, uint32_t, size_t>, uint32_t> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple(dev, fn, block, lds);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, (int)block, lds) != hipSuccess) {
        (void)hipGetLastError();
        nb = 0;
    }
    const uint32_t r = nb > 0 ? (uint32_t)nb : 1u;
    std::lock_guard<std::mutex> g(mu);
    cache[key] = r;
    return r;
}

int mul_snaps(bi_ctx* ctx, uint32_t** p, uint64_t* words) {
    const uint64_t w = mul_snap_words((uint32_t)ctx->num_cus);
    void* ptr;
    int rc = scratch(ctx, 5, w * 4, &ptr);
    if (rc) return rc;
    *p = (uint32_t*)ptr;
    *words = w;
    return BI_OK;
}

}  // namespace bi

extern "C" {

int bi_create(bi_ctx** out, int device) {
    if (!out) return BI_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return BI_ERR_NODEV;
    }
    if (device < 0 || device >= count) return BI_ERR_NODEV;
    bi_ctx* ctx = new (std::nothrow) bi_ctx();
    if (!ctx) return BI_ERR_NOMEM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete ctx;
        return BI_ERR_HIP;
    }
    tune_defaults(ctx);
    noalloc_init(ctx);
    *out = ctx;
    return BI_OK;
}

int bi_destroy(bi_ctx* ctx) {
    if (!ctx) return BI_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    destroy_extras(ctx);
    hostpipe_free(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < kScratchSlots; i++)
        if (ctx->scratch[i]) (void)hipFree(ctx->scratch[i]);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BI_OK;
}

const char* bi_strerror(int code) {
    switch (code) {
        case BI_OK: return "ok";
        case BI_ERR_ARG: return "invalid argument";
        case BI_ERR_SHAPE: return "unsupported size";
        case BI_ERR_HIP: return "HIP runtime error";
        case BI_ERR_NOMEM: return "out of device memory or over budget";
        case BI_ERR_DIVZERO: return "division by zero";
        case BI_ERR_BUSY: return "no free request slot";
        case BI_ERR_NODEV: return "no usable GPU";
        case BI_ERR_STATE: return "invalid state for this call";
        default: return "unknown error";
    }
}

const char* bi_last_hip_error(bi_ctx* ctx) { return ctx ? ctx->last_err : ""; }

int bi_sync(bi_ctx* ctx) {
    if (!ctx) return BI_ERR_ARG;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

void* bi_stream(bi_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int bi_dev_alloc(bi_ctx* ctx, void** ptr, size_t bytes) {
    if (!ctx || !ptr) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return BI_ERR_NOMEM;
    }
    return BI_OK;
}

int bi_dev_free(bi_ctx* ctx, void* ptr) {
    if (!ctx) return BI_ERR_ARG;
    if (!ptr) return BI_OK;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipFree(ptr);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_copy_h2d(bi_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return BI_ERR_ARG;
    if (!bytes) return BI_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_copy_d2h(bi_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return BI_ERR_ARG;
    if (!bytes) return BI_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_copy_d2d(bi_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return BI_ERR_ARG;
    if (!bytes) return BI_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_timer_start(bi_ctx* ctx) {
    if (!ctx) return BI_ERR_ARG;
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_timer_stop(bi_ctx* ctx, float* ms) {
    if (!ctx || !ms) return BI_ERR_ARG;
    hipError_t e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

// ---------------------------------------------------------------- add / mul
static int check_plain(bi_ctx* ctx, uint32_t m, const void* a, const void* b, const void* r, uint64_t n) {
    if (!ctx || m == 0) return BI_ERR_ARG;
    if (n && (!a || !b || !r)) return BI_ERR_ARG;
    if ((uint64_t)m * n > (1ull << 40)) return BI_ERR_SHAPE;
    return BI_OK;
}

int bi_add_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    int rc = check_plain(ctx, m, as, bs, rs, n);
    if (rc || n == 0) return rc;
    if ((rc = set_device(ctx))) return rc;
    const AddCfg& cfg = ctx->add_cfg_for(m);
    void* states = nullptr;
    const uint64_t sb = add_chunk_states(cfg, m, n);
    if (sb && (rc = scratch(ctx, 12, sb, &states))) return rc;
    hipError_t e = add_uniform(cfg, as, bs, rs, m, n, false, ctx->stream, nullptr, static_cast<uint8_t*>(states));
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_add_u64_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, const uint64_t* bs, uint64_t* rs) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return bi_add_u32_dev(ctx, 2 * m, n, (const uint32_t*)as, (const uint32_t*)bs, (uint32_t*)rs);
}

int bi_mul_u32_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    int rc = check_plain(ctx, m, as, bs, rs, n);
    if (rc || n == 0) return rc;
    if (m > kMulMaxDigits) return BI_ERR_SHAPE;
    if ((rc = set_device(ctx))) return rc;
    uint32_t* snaps;
    uint64_t snap_words;
    if ((rc = mul_snaps(ctx, &snaps, &snap_words))) return rc;
    hipError_t e = mul_uniform(ctx->mul_cfg, as, bs, rs, m, n, ctx->stream, snaps, snap_words, (uint32_t)ctx->num_cus);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_mul_u64_dev(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, const uint64_t* bs, uint64_t* rs) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return bi_mul_u32_dev(ctx, 2 * m, n, (const uint32_t*)as, (const uint32_t*)bs, (uint32_t*)rs);
}

// Host-pointer variants: the batch goes in and out through the context's
// transfer pipeline (bi_hostpipe.hip), device staging in scratch slots 0..2.
static int host_add(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    int rc = check_plain(ctx, m, as, bs, rs, n);
    if (rc || n == 0) return rc;
    if ((rc = set_device(ctx))) return rc;
    const AddCfg cfg = ctx->add_cfg_for(m);
    void* states = nullptr;
    const uint64_t sb = add_chunk_states(cfg, m, n);
    if (sb && (rc = scratch(ctx, 12, sb, &states))) return rc;
    return host_pipelined(ctx, m, n, as, bs, rs,
                          [&](hipStream_t st, const uint32_t* a, const uint32_t* b, uint32_t* r, uint64_t cnt) {
                              return add_uniform(cfg, a, b, r, m, cnt, false, st, nullptr,
                                                 static_cast<uint8_t*>(states));
                          });
}

static int host_mul(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    int rc = check_plain(ctx, m, as, bs, rs, n);
    if (rc || n == 0) return rc;
    if (m > kMulMaxDigits) return BI_ERR_SHAPE;
    if ((rc = set_device(ctx))) return rc;
    uint32_t* snaps;
    uint64_t snap_words;
    if ((rc = mul_snaps(ctx, &snaps, &snap_words))) return rc;
    const MulCfg cfg = ctx->mul_cfg;
    return host_pipelined(ctx, m, n, as, bs, rs,
                          [&](hipStream_t st, const uint32_t* a, const uint32_t* b, uint32_t* r, uint64_t cnt) {
                              return mul_uniform(cfg, a, b, r, m, cnt, st, snaps, snap_words, (uint32_t)ctx->num_cus);
                          });
}

int bi_add_u32(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    return host_add(ctx, m, n, as, bs, rs);
}
int bi_add_u64(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, const uint64_t* bs, uint64_t* rs) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return host_add(ctx, 2 * m, n, (const uint32_t*)as, (const uint32_t*)bs, (uint32_t*)rs);
}
int bi_mul_u32(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    return host_mul(ctx, m, n, as, bs, rs);
}
int bi_mul_u64(bi_ctx* ctx, uint32_t m, uint64_t n, const uint64_t* as, const uint64_t* bs, uint64_t* rs) {
    if (m > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return host_mul(ctx, 2 * m, n, (const uint32_t*)as, (const uint32_t*)bs, (uint32_t*)rs);
}

}  // extern "C"
