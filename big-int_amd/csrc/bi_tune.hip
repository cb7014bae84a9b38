// bi_tune.hip -- default kernel shapes (replaced by the autotuner's table).
#include "bi_ctx.h"

namespace bi {

void tune_defaults(bi_ctx* ctx) {
    for (int c = 0; c < 16; c++) {
        AddCfg a;
        a.span_digits = 1024;
        a.tiles_in_flight = 2;
        a.nontemporal = 1;
        ctx->add_cfg[c] = a;
    }
    ctx->mul_cfg.waves_per_block = 4;
}

void div_state_free(bi_ctx* ctx);

void destroy_extras(bi_ctx* ctx) { div_state_free(ctx); }

}  // namespace bi
