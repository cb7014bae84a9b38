// bi_noalloc.hip -- bi_mul_u32_dev_noalloc: the batched multiply of
// bi_mul_u32_dev (same kernels, same results) with its per-lane snapshot
// scratch in a static device array instead of context scratch, so that the
// call allocates nothing.  For callers that cannot allocate device memory:
// code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc.
//
// The scratch holds the multiply's per-lane snapshot: 12 words for each lane
// of the persistent grid (the first column group's sums recorded at a lane's
// switch, written before that same lane reads them, so nothing has to start
// at zero).  kNoAllocScratchWords covers 2048 lanes on each of 256 CUs, a full
// MI355X; on a device with more CUs the persistent grid is cut to what the
// array holds (same results, fewer resident workgroups).  Its size does not
// grow with n or m, so no batch is too large for it; m above 4096 digits is
// rejected as by bi_mul_u32_dev.
//
// Each GPU holds its own copy of the array, at its own address, until the
// process ends.  Every call on one GPU shares that copy: calls that overlap
// in time on one GPU (two streams, two contexts, two host threads) race on
// it, and must be ordered by the caller.
#include <string.h>

#include <mutex>

#include "../../include/bigint.h"
#include "bi_ctx.h"

namespace {

constexpr uint64_t kNoAllocScratchWords = 256ull * 2048 * 12;   // 24 MiB
constexpr int kNoAllocMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {

std::mutex g_mu;
uint32_t* g_ptr[kNoAllocMaxDevices] = {};   // this device's copy of `scratch`

void set_err(bi_ctx* ctx, const char* msg) {
    strncpy(ctx->last_err, msg, sizeof(ctx->last_err) - 1);
    ctx->last_err[sizeof(ctx->last_err) - 1] = 0;
}

}  // namespace

namespace bi {

// Resolves the device address of `scratch` on ctx->device (which the caller,
// bi_create, has made current), once per device, and keeps it in the context.
// A failure is recorded, not returned: only the allocation-free entry point
// depends on it, and it reports the error.
void noalloc_init(bi_ctx* ctx) {
    ctx->noalloc_scratch = nullptr;
    if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[ctx->device]) {
        void* p = nullptr;
        // ::scratch is the array (bi::scratch is the context allocator)
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return;
        }
        g_ptr[ctx->device] = static_cast<uint32_t*>(p);
    }
    ctx->noalloc_scratch = g_ptr[ctx->device];
}

}  // namespace bi

extern "C" {

int bi_mul_u32_dev_noalloc(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs,
                           uint32_t* rs) {
    if (!ctx || m == 0) return BI_ERR_ARG;
    if (n && (!as || !bs || !rs)) return BI_ERR_ARG;
    if ((uint64_t)m * n > (1ull << 40)) return BI_ERR_SHAPE;
    if (n == 0) return BI_OK;
    if (m > bi::kMulMaxDigits) {
        set_err(ctx, "bi_mul_u32_dev_noalloc: m exceeds 4096 digits");
        return BI_ERR_SHAPE;
    }
    if (!ctx->noalloc_scratch) {
        set_err(ctx, "bi_mul_u32_dev_noalloc: static scratch address unavailable on this device");
        return BI_ERR_STATE;
    }
    int rc = bi::set_device(ctx);
    if (rc) return rc;
    const uint64_t need = bi::mul_snap_words((uint32_t)ctx->num_cus);
    const uint64_t words = need < kNoAllocScratchWords ? need : kNoAllocScratchWords;
    const hipError_t e = bi::mul_uniform(ctx->mul_cfg, as, bs, rs, m, n, ctx->stream, ctx->noalloc_scratch, words,
                                         (uint32_t)ctx->num_cus);
    return e == hipSuccess ? BI_OK : bi::hip_fail(ctx, e);
}

}  // extern "C"
