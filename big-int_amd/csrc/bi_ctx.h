// bi_ctx.h -- the context behind the opaque bi_ctx handle (internal).
#pragma once
#include "bi_kernels.h"

namespace bi {

constexpr int kScratchSlots = 8;

}  // namespace bi

struct bi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* scratch[bi::kScratchSlots] = {};
    size_t scratch_bytes[bi::kScratchSlots] = {};
    char last_err[256] = {0};
    // tuned shapes: add/sub per size class (log2 of digits, 0..15), multiply
    bi::AddCfg add_cfg[16];
    bi::MulCfg mul_cfg;
    const bi::AddCfg& add_cfg_for(uint64_t m) const {
        int c = 0;
        while (c < 15 && (1ull << (c + 1)) <= m) c++;
        return add_cfg[c];
    }
    void* extras = nullptr;  // division / serving state (owned, released by destroy_extras)
};

namespace bi {
int hip_fail(bi_ctx* ctx, hipError_t e);
int set_device(bi_ctx* ctx);
int scratch(bi_ctx* ctx, int slot, size_t bytes, void** out);
void tune_defaults(bi_ctx* ctx);
void destroy_extras(bi_ctx* ctx);
}  // namespace bi
