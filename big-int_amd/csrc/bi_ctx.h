// bi_ctx.h -- the context behind the opaque bi_ctx handle (internal).
#pragma once
#include <functional>

#include "bi_kernels.h"

namespace bi {

constexpr int kScratchSlots = 13;   // slot 12: chunk states of the long-instance add

}  // namespace bi

struct bi_plan;

struct bi_ctx {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* scratch[bi::kScratchSlots] = {};
    size_t scratch_bytes[bi::kScratchSlots] = {};
    char last_err[256] = {0};
    // tuned shapes: add and |a-b| per size class (log2 of digits, 0..15),
    // multiply, ragged add and ragged multiply
    bi::AddCfg add_cfg[16];
    bi::AddCfg sub_cfg[16];    // |a-b| with the sign decided by each span wave
    bi::MulCfg mul_cfg;
    bi::AddCfg csr_add_cfg;    // ragged add (span digits of whole instances)
    uint32_t csr_lanes = 256;  // lanes per ragged-multiply workgroup
    uint32_t csr_rows = 8;     // row digits per block step of the ragged multiply
    uint32_t div_one_max = 4096;  // largest batch the one-launch division takes (0: never)
    static int size_class(uint64_t m) {
        int c = 0;
        while (c < 15 && (1ull << (c + 1)) <= m) c++;
        return c;
    }
    const bi::AddCfg& add_cfg_for(uint64_t m) const { return add_cfg[size_class(m)]; }
    const bi::AddCfg& sub_cfg_for(uint64_t m) const { return sub_cfg[size_class(m)]; }
    void* extras = nullptr;  // division / serving state (owned, released by destroy_extras)
    uint32_t* noalloc_scratch = nullptr;  // this device's static multiply scratch (bi_noalloc.hip)
    void* hostpipe = nullptr;  // host-pointer transfer pipeline (bi_hostpipe.hip), created on first use
};

namespace bi {
int hip_fail(bi_ctx* ctx, hipError_t e);
int set_device(bi_ctx* ctx);
int scratch(bi_ctx* ctx, int slot, size_t bytes, void** out);
// per-lane snapshot scratch of the persistent multiply kernels (slot 5)
int mul_snaps(bi_ctx* ctx, uint32_t** p, uint64_t* words);
void tune_defaults(bi_ctx* ctx);
// enqueues the kernels of `count` instances (device pointers) on a stream
typedef std::function<hipError_t(hipStream_t, const uint32_t*, const uint32_t*, uint32_t*, uint64_t)> HostLaunch;
// host arrays in and out through pinned staging, in chunks on two streams
// (chunk_hint: instances per chunk, 0 = about 8 MiB of each array)
int host_pipelined(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs,
                   const HostLaunch& launch, uint64_t chunk_hint = 0);
void hostpipe_free(bi_ctx* ctx);
// resolves ctx->noalloc_scratch on ctx->device (current); bi_noalloc.hip
void noalloc_init(bi_ctx* ctx);
void destroy_extras(bi_ctx* ctx);
int div_reserve(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, uint32_t P);
int divmod(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs, uint32_t* qs,
           uint32_t* rs, bool graph);
// the one-workgroup division (bi_divone.hip): shapes whose LDS layout fits
// one workgroup, and its launch (bad[i] = 1 for a zero divisor)
bool divone_fits(uint32_t mu, uint32_t mv);
bool divone_wanted(uint64_t n, uint32_t mu, uint32_t mv, uint32_t n_max);
hipError_t divone_launch(uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs, uint32_t* qs,
                         uint32_t* rs, uint8_t* bad, hipStream_t st);
int csr_plan_rebuild(bi_ctx* ctx, bi_plan** pp, uint64_t n, const int64_t* offs, int mode);
int csr_plan_reserve(bi_ctx* ctx, bi_plan** pp, size_t teams);
void csr_plan_free(bi_plan* p);
int addmul_csr_split(bi_ctx* ctx, const bi_plan* plan, uint64_t n_add, uint64_t add_total, const int64_t* offs,
                     const uint32_t* as, const uint32_t* bs, uint32_t* rs_add, uint32_t* rs_mul);
// transform products (bi_ntt.hip)
struct NttPlan {
    int lg;   // log2 points, -1 when the product does not fit
    int w;    // chunk bits, 16 or 25
};
NttPlan ntt_plan(uint32_t ma, uint32_t mb);
int ntt_prepare(bi_ctx* ctx, uint64_t n, const uint64_t** wf, const uint64_t** wi, uint64_t** stash);
// tb_save / tb_load (2^13 points, 25-bit chunks only): n * 2^13 words of the
// second operand's transform, stored for a later product with the same
// second operand, or taken from an earlier one instead of being computed
hipError_t ntt_launch(bi_ctx* ctx, uint64_t n, uint32_t ma, uint32_t mb, const uint32_t* as, uint64_t sa,
                      const uint32_t* bs, uint64_t sb, uint32_t* rs, uint64_t sr, uint32_t rlo, uint32_t rlen,
                      bool neg, hipStream_t st, bool cyclic = false, uint64_t* tb_save = nullptr,
                      const uint64_t* tb_load = nullptr);
// Digits of a*b mod (B^kCyclicDigits - 1) come out of the cyclic mode
// (2^13 points of 25-bit chunks, 25 * 2^13 = 32 * 6400 bits): digits
// [0, kCyclicDigits + 3) of the wrapped convolution's value, whose top
// digits still have to be folded back (ma, mb <= kCyclicDigits).
constexpr uint32_t kCyclicDigits = 6400;
}  // namespace bi
