// bi_tune.hip -- kernel shape table and the autotuner that fills it.
//
// Tunables (none changes any result, only the launch shape):
//   add/sub  per size class (log2 digits): digits per wave span, 256-digit
//            tiles whose loads are issued together, nontemporal streaming
//   mul      limb width of the uniform multiply's convolution (32-bit digits,
//            or 27-bit limbs cut from them for m <= 864: one v_mad_u64_u32 per
//            limb product, no overflow counts), workgroup size (waves) for
//            m <= 512, and digits of the row operand per block step (4 or 8
//            rows x 4 columns = 16 or 32 multiply-adds per lane per step)
//   ragged   lanes per workgroup bundle of the ragged multiply
// u32 against u64 limbs is not a tunable: a u64-limb batch is the same bytes
// as a u32-digit batch with twice the digits, and v_mad_u64_u32 is a
// 32x32->64 multiply (a 64-bit limb product would be four of them).
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/bigint.h"
#include "bi_ctx.h"

namespace bi {

void tune_defaults(bi_ctx* ctx) {
    for (int c = 0; c < 16; c++) {
        AddCfg a;
        a.span_digits = 1024;
        a.tiles_in_flight = 1;   // measured best untuned: 131072 x 512 uniform, 2^17 ragged
        a.nontemporal = 1;
        ctx->add_cfg[c] = a;
    }
    ctx->mul_cfg.waves_per_block = 4;
    ctx->mul_cfg.rows = 8;
    ctx->csr_lanes = 256;
}

void div_state_free(bi_ctx* ctx);

void destroy_extras(bi_ctx* ctx) { div_state_free(ctx); }

static int size_class(uint64_t m) {
    int c = 0;
    while (c < 15 && (1ull << (c + 1)) <= m) c++;
    return c;
}

}  // namespace bi

using namespace bi;

namespace {

struct Bufs {
    uint32_t *a = nullptr, *b = nullptr, *r = nullptr;
    ~Bufs() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (r) (void)hipFree(r);
    }
};

__global__ void k_fill(uint32_t* p, uint64_t n, uint32_t seed) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint64_t x = (i + 1) * 0x9E3779B97F4A7C15ull ^ seed;
        x ^= x >> 31;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 29;
        p[i] = (uint32_t)x;
    }
}

template <class F>
float time_ms(bi_ctx* ctx, F fn, int reps) {
    fn();
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    for (int i = 0; i < reps; i++) fn();
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    (void)hipEventSynchronize(ctx->ev1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    return ms / reps;
}

}  // namespace

extern "C" {

// Sweeps the add shape for integers of m digits (batch of ~bytes per operand)
// and the multiply workgroup size; stores the best shapes in the context.
int bi_tune_run(bi_ctx* ctx, uint32_t m, uint64_t bytes_per_operand, int what, float* best_ms) {
    if (!ctx || m == 0) return BI_ERR_ARG;
    int rc = set_device(ctx);
    if (rc) return rc;
    const uint64_t n = std::max<uint64_t>(1, bytes_per_operand / (4ull * m));
    const uint64_t words = n * m;
    Bufs B;
    hipError_t e = hipMalloc(&B.a, words * 4);
    if (e == hipSuccess) e = hipMalloc(&B.b, words * 4);
    if (e == hipSuccess) e = hipMalloc(&B.r, words * 4);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BI_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, ctx->stream, B.a, words, 1u);
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, ctx->stream, B.b, words, 2u);
    float best = 1e30f;
    if (what == 0) {   // add
        AddCfg bestc = ctx->add_cfg[size_class(m)];
        const uint32_t spans[] = {256, 512, 1024, 2048, 4096, 8192};
        const uint32_t tiles[] = {1, 2, 4};
        for (uint32_t sp : spans) {
            for (uint32_t t : tiles) {
                for (uint32_t nt = 0; nt < 2; nt++) {
                    AddCfg c;
                    c.span_digits = sp;
                    c.tiles_in_flight = t;
                    c.nontemporal = nt;
                    const float ms = time_ms(
                        ctx, [&]() { (void)add_uniform(c, B.a, B.b, B.r, m, n, false, ctx->stream); }, 10);
                    if (ms < best) {
                        best = ms;
                        bestc = c;
                    }
                }
            }
        }
        ctx->add_cfg[size_class(m)] = bestc;
    } else if (what == 1) {   // uniform multiply
        if (m > kMulMaxDigits) return BI_ERR_SHAPE;
        MulCfg bestc = ctx->mul_cfg;
        uint32_t* snaps;
        uint64_t snap_words;
        if ((rc = mul_snaps(ctx, &snaps, &snap_words))) return rc;
        const uint32_t wpbs[] = {1, 2, 4, 8};
        const uint32_t rows[] = {4, 8};
        if (mul27_eligible(m)) {   // the limb convolution shapes itself
            MulCfg c = ctx->mul_cfg;
            c.limb_bits = 27;
            const float ms =
                time_ms(ctx, [&]() { (void)mul_uniform(c, B.a, B.b, B.r, m, n, ctx->stream, snaps, snap_words,
                                                        (uint32_t)ctx->num_cus); },
                        3);
            if (ms < best) {
                best = ms;
                bestc = c;
            }
        }
        for (uint32_t rr : rows) {
            for (uint32_t w : wpbs) {
                MulCfg c;
                c.waves_per_block = w;
                c.rows = rr;
                c.limb_bits = 32;
                const float ms =
                    time_ms(ctx, [&]() { (void)mul_uniform(c, B.a, B.b, B.r, m, n, ctx->stream, snaps, snap_words,
                                                            (uint32_t)ctx->num_cus); },
                            3);
                if (ms < best) {
                    best = ms;
                    bestc = c;
                }
            }
        }
        ctx->mul_cfg = bestc;
    } else {
        return BI_ERR_ARG;
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if (best_ms) *best_ms = best;
    return BI_OK;
}

// Sets the ragged multiply's lanes per workgroup (64..1024, multiple of 64).
int bi_tune_set_csr_lanes(bi_ctx* ctx, uint32_t lanes) {
    if (!ctx || lanes < 64 || lanes > 1024 || (lanes & 63)) return BI_ERR_ARG;
    ctx->csr_lanes = lanes;
    return BI_OK;
}

// Sets one add shape explicitly (size class of m digits).
int bi_tune_set_add(bi_ctx* ctx, uint32_t m, uint32_t span_digits, uint32_t tiles_in_flight, uint32_t nontemporal) {
    if (!ctx || m == 0 || span_digits < 1 || (tiles_in_flight != 1 && tiles_in_flight != 2 && tiles_in_flight != 4))
        return BI_ERR_ARG;
    AddCfg c;
    c.span_digits = span_digits;
    c.tiles_in_flight = tiles_in_flight;
    c.nontemporal = nontemporal ? 1 : 0;
    ctx->add_cfg[size_class(m)] = c;
    return BI_OK;
}

int bi_tune_set_mul(bi_ctx* ctx, uint32_t waves_per_block, uint32_t rows) {
    if (!ctx || waves_per_block < 1 || waves_per_block > 16 || (rows != 4 && rows != 8)) return BI_ERR_ARG;
    ctx->mul_cfg.waves_per_block = waves_per_block;
    ctx->mul_cfg.rows = rows;
    return BI_OK;
}

// Chooses the multiply's convolution: 27-bit limbs (m <= 864) or 32-bit digits.
int bi_tune_set_mul_limb_bits(bi_ctx* ctx, uint32_t bits) {
    if (!ctx || (bits != 27 && bits != 32)) return BI_ERR_ARG;
    ctx->mul_cfg.limb_bits = bits;
    return BI_OK;
}

int bi_tune_get_mul_limb_bits(bi_ctx* ctx, uint32_t* bits) {
    if (!ctx || !bits) return BI_ERR_ARG;
    *bits = ctx->mul_cfg.limb_bits;
    return BI_OK;
}

// Text table: "add <class> <span> <tiles> <nt>", "mul <wpb> <rows>", "mullimb <bits>",
// "csr <lanes>".
int bi_tune_save(bi_ctx* ctx, const char* path) {
    if (!ctx || !path) return BI_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return BI_ERR_ARG;
    fprintf(f, "# libbigint tuned shapes\n");
    for (int c = 0; c < 16; c++)
        fprintf(f, "add %d %u %u %u\n", c, ctx->add_cfg[c].span_digits, ctx->add_cfg[c].tiles_in_flight,
                ctx->add_cfg[c].nontemporal);
    fprintf(f, "mul %u %u\n", ctx->mul_cfg.waves_per_block, ctx->mul_cfg.rows);
    fprintf(f, "mullimb %u\n", ctx->mul_cfg.limb_bits);
    fprintf(f, "csr %u\n", ctx->csr_lanes);
    fclose(f);
    return BI_OK;
}

int bi_tune_load(bi_ctx* ctx, const char* path) {
    if (!ctx || !path) return BI_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f) return BI_ERR_ARG;
    char line[256];
    int bad = 0;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        int c;
        unsigned a, b, d;
        if (sscanf(line, "add %d %u %u %u", &c, &a, &b, &d) == 4 && c >= 0 && c < 16 && a >= 1 &&
            (b == 1 || b == 2 || b == 4)) {
            ctx->add_cfg[c].span_digits = a;
            ctx->add_cfg[c].tiles_in_flight = b;
            ctx->add_cfg[c].nontemporal = d ? 1 : 0;
        } else if (sscanf(line, "mul %u %u", &a, &b) == 2 && a >= 1 && a <= 16 && (b == 4 || b == 8)) {
            ctx->mul_cfg.waves_per_block = a;
            ctx->mul_cfg.rows = b;
        } else if (sscanf(line, "mullimb %u", &a) == 1 && (a == 27 || a == 32)) {
            ctx->mul_cfg.limb_bits = a;
        } else if (sscanf(line, "csr %u", &a) == 1 && a >= 64 && a <= 1024 && !(a & 63)) {
            ctx->csr_lanes = a;
        } else {
            bad = 1;
        }
    }
    fclose(f);
    return bad ? BI_ERR_ARG : BI_OK;
}

int bi_tune_get(bi_ctx* ctx, uint32_t m, uint32_t* span_digits, uint32_t* tiles_in_flight, uint32_t* nontemporal,
                uint32_t* mul_waves, uint32_t* mul_rows, uint32_t* csr_lanes) {
    if (!ctx) return BI_ERR_ARG;
    const AddCfg& c = ctx->add_cfg_for(m ? m : 1);
    if (span_digits) *span_digits = c.span_digits;
    if (tiles_in_flight) *tiles_in_flight = c.tiles_in_flight;
    if (nontemporal) *nontemporal = c.nontemporal;
    if (mul_waves) *mul_waves = ctx->mul_cfg.waves_per_block;
    if (mul_rows) *mul_rows = ctx->mul_cfg.rows;
    if (csr_lanes) *csr_lanes = ctx->csr_lanes;
    return BI_OK;
}

}  // extern "C"
