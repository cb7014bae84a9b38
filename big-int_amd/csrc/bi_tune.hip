// bi_tune.hip -- kernel shape table and the autotuner that fills it.
//
// Tunables (none changes any result, only the launch shape):
//   add      per size class (log2 digits): digits per wave span, 256-digit
//            tiles whose loads are issued together, nontemporal streaming
//   sub      the same three for |a-b| (each span wave first decides a < b
//            for its own instances, so its best span differs from the add's)
//   mul      limb width of the uniform multiply's convolution (32-bit digits,
//            or 27-bit limbs cut from them for m <= 864: one v_mad_u64_u32 per
//            limb product, no overflow counts), workgroup size (waves) for
//            m <= 512, and digits of the row operand per block step (4 or 8
//            rows x 4 columns = 16 or 32 multiply-adds per lane per step)
//   ragged   the ragged add's span and tiles in flight; lanes per workgroup
//            bundle and rows per block step of the ragged multiply
// u32 against u64 limbs is not a tunable: a u64-limb batch is the same bytes
// as a u32-digit batch with twice the digits, and v_mad_u64_u32 is a
// 32x32->64 multiply (a 64-bit limb product would be four of them).
//
// Timing.  Candidates are timed in interleaved rounds (every round runs each
// candidate once: one untimed launch, then a few timed ones, starting at a
// different candidate each round), a candidate's score is the median of its
// rounds, and the shape in place is kept unless the best candidate beats it
// by kKeepMargin.  A single pass that keeps the minimum lets launch-to-launch
// noise of a few percent pick the winner, and on a fresh GPU it picked a
// shape 3 % slower than the default for 131072 x 512 digits.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
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
        ctx->sub_cfg[c] = a;
        // 2048 digits and up: 1024-digit chunks with 2 tiles in flight (2^27-digit
        // batches of 2048 / 4096 / 8192 digits: 6.23 / 6.29 / 6.35 TB/s, one tile
        // 6.07 / 6.16 / 6.25)
        if (c >= 11) a.tiles_in_flight = 2;
        ctx->add_cfg[c] = a;
    }
    // the ragged add: 2048-digit chunks, 4 tiles in flight (2^17 instances of
    // 8..2048 digits: 101.7 us; 1024 digits 106.5, 1536 105.0, 3072 109.8)
    ctx->csr_add_cfg = AddCfg();
    ctx->csr_add_cfg.span_digits = 2048;
    ctx->csr_add_cfg.tiles_in_flight = 4;
    ctx->mul_cfg.waves_per_block = 4;
    ctx->mul_cfg.rows = 8;
    ctx->csr_lanes = 256;
    ctx->csr_rows = 8;
}

void div_state_free(bi_ctx* ctx);

void destroy_extras(bi_ctx* ctx) { div_state_free(ctx); }

}  // namespace bi

using namespace bi;

namespace {

constexpr float kKeepMargin = 0.02f;   // a new shape must be 2 % faster
constexpr int kRounds = 5;

struct Bufs {
    uint32_t *a = nullptr, *b = nullptr, *r = nullptr, *r2 = nullptr;
    uint8_t* order = nullptr;
    uint8_t* states = nullptr;
    int64_t* off = nullptr;
    ~Bufs() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (r) (void)hipFree(r);
        if (r2) (void)hipFree(r2);
        if (order) (void)hipFree(order);
        if (states) (void)hipFree(states);
        if (off) (void)hipFree(off);
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

void fill(bi_ctx* ctx, uint32_t* p, uint64_t words, uint32_t seed) {
    if (words) hipLaunchKernelGGL(k_fill, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, ctx->stream, p, words, seed);
}

float median(std::vector<float> v) {
    std::sort(v.begin(), v.end());
    const size_t k = v.size();
    return k % 2 ? v[k / 2] : 0.5f * (v[k / 2 - 1] + v[k / 2]);
}

// Median ms per launch of each candidate (run(c) launches candidate c on the
// context stream), over kRounds interleaved rounds of `launches` timed calls.
template <class Run>
std::vector<float> time_candidates(bi_ctx* ctx, size_t ncand, Run run, int launches) {
    std::vector<std::vector<float>> t(ncand);
    for (int r = 0; r < kRounds; r++) {
        for (size_t k = 0; k < ncand; k++) {
            const size_t c = (k + (size_t)r * 7) % ncand;
            run(c);
            (void)hipEventRecord(ctx->ev0, ctx->stream);
            for (int i = 0; i < launches; i++) run(c);
            (void)hipEventRecord(ctx->ev1, ctx->stream);
            (void)hipEventSynchronize(ctx->ev1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
            t[c].push_back(ms / launches);
        }
    }
    std::vector<float> med(ncand);
    for (size_t c = 0; c < ncand; c++) med[c] = median(t[c]);
    return med;
}

// Index of the candidate to keep: the fastest, unless it beats candidate
// `cur` (the shape in place) by less than kKeepMargin.
size_t choose(const std::vector<float>& med, size_t cur, float* best_ms) {
    size_t b = 0;
    for (size_t c = 1; c < med.size(); c++)
        if (med[c] < med[b]) b = c;
    if (cur < med.size() && !(med[b] < (1.0f - kKeepMargin) * med[cur])) b = cur;
    if (best_ms) *best_ms = med[b];
    return b;
}

bool same(const AddCfg& x, const AddCfg& y) {
    return x.span_digits == y.span_digits && x.tiles_in_flight == y.tiles_in_flight && x.nontemporal == y.nontemporal;
}

// What a span length does for instances of m digits: a run of whole
// instances (positive: instances per span) or, for instances longer than
// the span, chunks of the span's whole tiles (negative: tiles per chunk).
int64_t span_key(uint32_t span, uint32_t m) {
    if ((span & ~255u) >= 256 && m > (span & ~255u)) return -(int64_t)(span / 256);
    return std::max<uint32_t>(1, span / m);
}

// Span / tiles / nontemporal candidates for a uniform size m (spans that give
// the same whole-instance count are one candidate); `cur` first.
std::vector<AddCfg> add_candidates(const AddCfg& cur, uint32_t m, bool uniform) {
    std::vector<AddCfg> v{cur};
    const uint32_t spans[] = {256, 512, 768, 1024, 1536, 2048, 4096, 8192};
    const uint32_t tiles[] = {1, 2, 4, 8};   // 8: the ragged add's fixed chunks only
    for (uint32_t sp : spans)
        for (uint32_t t : tiles)
            for (uint32_t nt = 0; nt < 2; nt++) {
                if (t == 8 && uniform) continue;
                AddCfg c;
                c.span_digits = sp;
                c.tiles_in_flight = t;
                c.nontemporal = nt;
                bool dup = false;
                for (const AddCfg& x : v) {
                    const bool eq_span = uniform ? span_key(x.span_digits, m) == span_key(sp, m)
                                                 : x.span_digits == sp;
                    dup |= eq_span && x.tiles_in_flight == t && x.nontemporal == nt;
                }
                if (!dup) v.push_back(c);
            }
    return v;
}

// A ragged sample: digit counts log-uniform in [8, maxd], about `digits`
// digits in all (host offsets).
std::vector<int64_t> ragged_sample(uint32_t maxd, uint64_t digits) {
    std::vector<int64_t> off{0};
    uint64_t x = 0x243F6A8885A308D3ull;
    const double lo = log(8.0), hi = log((double)std::max<uint32_t>(maxd, 8));
    while ((uint64_t)off.back() < digits && off.size() < (1u << 24)) {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        const uint32_t d = std::min<uint32_t>(maxd, (uint32_t)llround(exp(lo + u * (hi - lo))));
        off.push_back(off.back() + d);
    }
    return off;
}

}  // namespace

extern "C" {

// what 0: add shape for the size class of m (n x m batch of ~bytes per
//         operand); 1: uniform multiply; 2: ragged add and ragged multiply on
//         a sample of digit counts log-uniform in [8, m] (m <= 4096) of ~bytes
//         per operand; 3: |a-b| shape for the size class of m.
int bi_tune_run(bi_ctx* ctx, uint32_t m, uint64_t bytes_per_operand, int what, float* best_ms) {
    if (!ctx || m == 0 || what < 0 || what > 3) return BI_ERR_ARG;
    if ((what == 1 || what == 2) && m > kMulMaxDigits) return BI_ERR_SHAPE;
    int rc = set_device(ctx);
    if (rc) return rc;
    Bufs B;
    std::vector<int64_t> hoff;
    uint64_t n = std::max<uint64_t>(1, bytes_per_operand / (4ull * m));
    uint64_t words = n * m;
    if (what == 2) {
        hoff = ragged_sample(m, std::max<uint64_t>(bytes_per_operand / 4, 4096));
        n = hoff.size() - 1;
        words = (uint64_t)hoff.back();
    }
    hipError_t e = hipMalloc(&B.a, words * 4);
    if (e == hipSuccess) e = hipMalloc(&B.b, words * 4);
    if (e == hipSuccess) e = hipMalloc(&B.r, words * 4);
    if (e == hipSuccess && what == 2) e = hipMalloc(&B.r2, words * 4);
    if (e == hipSuccess && what == 2) e = hipMalloc(&B.off, hoff.size() * 8);
    if (e == hipSuccess && what == 3) e = hipMalloc(&B.order, n);
    if (e == hipSuccess && what == 0) e = hipMalloc(&B.states, n * ((m + 255) / 256));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BI_ERR_NOMEM;
    }
    if (what == 2) {
        e = hipMemcpy(B.off, hoff.data(), hoff.size() * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(ctx, e);
    }
    fill(ctx, B.a, words, 1u);
    fill(ctx, B.b, words, 2u);
    float best = 0;
    if (what == 0 || what == 3) {   // add, |a-b|
        const bool sub = what == 3;
        AddCfg& slot = sub ? ctx->sub_cfg[bi_ctx::size_class(m)] : ctx->add_cfg[bi_ctx::size_class(m)];
        const std::vector<AddCfg> cand = add_candidates(slot, m, true);
        const std::vector<float> med = time_candidates(
            ctx, cand.size(),
            [&](size_t c) { (void)add_uniform(cand[c], B.a, B.b, B.r, m, n, sub, ctx->stream, B.order, B.states); },
            10);
        slot = cand[choose(med, 0, &best)];
    } else if (what == 1) {   // uniform multiply
        uint32_t* snaps;
        uint64_t snap_words;
        if ((rc = mul_snaps(ctx, &snaps, &snap_words))) return rc;
        std::vector<MulCfg> cand{ctx->mul_cfg};
        if (mul27_eligible(m)) {   // the limb convolution shapes itself
            MulCfg c = ctx->mul_cfg;
            c.limb_bits = 27;
            if (ctx->mul_cfg.limb_bits != 27) cand.push_back(c);
        }
        for (uint32_t rr : {4u, 8u})
            for (uint32_t w : {1u, 2u, 4u, 8u}) {
                MulCfg c;
                c.waves_per_block = w;
                c.rows = rr;
                c.limb_bits = 32;
                const MulCfg& k = cand[0];
                if (!(k.limb_bits == 32 && k.waves_per_block == w && k.rows == rr)) cand.push_back(c);
            }
        const std::vector<float> med = time_candidates(
            ctx, cand.size(),
            [&](size_t c) {
                (void)mul_uniform(cand[c], B.a, B.b, B.r, m, n, ctx->stream, snaps, snap_words, (uint32_t)ctx->num_cus);
            },
            3);
        ctx->mul_cfg = cand[choose(med, 0, &best)];
    } else {   // ragged add, then ragged multiply
        const uint64_t total = words;
        const std::vector<AddCfg> acand = add_candidates(ctx->csr_add_cfg, 0, false);
        std::vector<AddCfg> ac;
        for (const AddCfg& c : acand)   // streaming loads only: the ragged stream is read once
            if (c.nontemporal && c.span_digits >= 512 && c.span_digits <= 4096) ac.push_back(c);
        if (!same(ac[0], ctx->csr_add_cfg)) ac.insert(ac.begin(), ctx->csr_add_cfg);
        const std::vector<float> amed = time_candidates(
            ctx, ac.size(), [&](size_t c) { (void)add_csr(ac[c], B.a, B.b, B.r, B.off, n, total, false, ctx->stream); },
            10);
        float add_best = 0;
        ctx->csr_add_cfg = ac[choose(amed, 0, &add_best)];
        // ragged multiply: one plan per lane budget, rows 4 / 8
        struct MC {
            uint32_t lanes, rows;
        };
        std::vector<MC> mc{{ctx->csr_lanes, ctx->csr_rows}};
        for (uint32_t l : {64u, 128u, 256u, 512u, 1024u})
            for (uint32_t rr : {4u, 8u})
                if (!(l == ctx->csr_lanes && rr == ctx->csr_rows)) mc.push_back({l, rr});
        std::vector<bi_plan*> plans(mc.size(), nullptr);
        for (size_t c = 0; c < mc.size() && rc == BI_OK; c++) {
            for (size_t k = 0; k < c && !plans[c]; k++)
                if (mc[k].lanes == mc[c].lanes) plans[c] = plans[k];
            if (!plans[c]) rc = bi_csr_plan_create(ctx, n, hoff.data(), 0, mc[c].lanes, &plans[c]);
        }
        const uint32_t rows0 = ctx->csr_rows;
        std::vector<float> mmed;
        if (rc == BI_OK)
            mmed = time_candidates(
                ctx, mc.size(),
                [&](size_t c) {
                    ctx->csr_rows = mc[c].rows;
                    (void)bi_mul_csr_u32_dev(ctx, plans[c], B.off, nullptr, B.a, B.b, B.r2);
                },
                3);
        ctx->csr_rows = rows0;
        for (size_t c = 0; c < mc.size(); c++) {
            bool dup = false;
            for (size_t k = 0; k < c; k++) dup |= plans[k] == plans[c];
            if (plans[c] && !dup) (void)bi_csr_plan_destroy(plans[c]);
        }
        if (rc) return rc;
        const size_t k = choose(mmed, 0, &best);
        ctx->csr_lanes = mc[k].lanes;
        ctx->csr_rows = mc[k].rows;
        best += add_best;
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if (best_ms) *best_ms = best;
    return BI_OK;
}

static bool add_shape_ok(uint32_t span_digits, uint32_t tiles_in_flight) {
    return span_digits >= 1 && (tiles_in_flight == 1 || tiles_in_flight == 2 || tiles_in_flight == 4 ||
                                tiles_in_flight == 8);
}

static AddCfg add_shape(uint32_t span_digits, uint32_t tiles_in_flight, uint32_t nontemporal) {
    AddCfg c;
    c.span_digits = span_digits;
    c.tiles_in_flight = tiles_in_flight;
    c.nontemporal = nontemporal ? 1 : 0;
    return c;
}

// Sets the ragged multiply's lanes per workgroup (64..1024, multiple of 64).
int bi_tune_set_csr_lanes(bi_ctx* ctx, uint32_t lanes) {
    if (!ctx || lanes < 64 || lanes > 1024 || (lanes & 63)) return BI_ERR_ARG;
    ctx->csr_lanes = lanes;
    return BI_OK;
}

// Sets one add shape explicitly (size class of m digits).
int bi_tune_set_add(bi_ctx* ctx, uint32_t m, uint32_t span_digits, uint32_t tiles_in_flight, uint32_t nontemporal) {
    if (!ctx || m == 0 || !add_shape_ok(span_digits, tiles_in_flight)) return BI_ERR_ARG;
    ctx->add_cfg[bi_ctx::size_class(m)] = add_shape(span_digits, tiles_in_flight, nontemporal);
    return BI_OK;
}

// Sets one |a-b| shape explicitly (size class of m digits).
int bi_tune_set_sub(bi_ctx* ctx, uint32_t m, uint32_t span_digits, uint32_t tiles_in_flight, uint32_t nontemporal) {
    if (!ctx || m == 0 || !add_shape_ok(span_digits, tiles_in_flight)) return BI_ERR_ARG;
    ctx->sub_cfg[bi_ctx::size_class(m)] = add_shape(span_digits, tiles_in_flight, nontemporal);
    return BI_OK;
}

// Sets the ragged shapes: add span (digits of whole instances), tiles in
// flight, nontemporal; multiply lanes per bundle and rows per block step.
int bi_tune_set_csr(bi_ctx* ctx, uint32_t add_span_digits, uint32_t add_tiles_in_flight, uint32_t add_nontemporal,
                    uint32_t mul_lanes, uint32_t mul_rows) {
    if (!ctx || !add_shape_ok(add_span_digits, add_tiles_in_flight) || mul_lanes < 64 || mul_lanes > 1024 ||
        (mul_lanes & 63) || (mul_rows != 4 && mul_rows != 8))
        return BI_ERR_ARG;
    ctx->csr_add_cfg = add_shape(add_span_digits, add_tiles_in_flight, add_nontemporal);
    ctx->csr_lanes = mul_lanes;
    ctx->csr_rows = mul_rows;
    return BI_OK;
}

int bi_tune_set_mul(bi_ctx* ctx, uint32_t waves_per_block, uint32_t rows) {
    if (!ctx || waves_per_block < 1 || waves_per_block > 16 || (rows != 4 && rows != 8)) return BI_ERR_ARG;
    ctx->mul_cfg.waves_per_block = waves_per_block;
    ctx->mul_cfg.rows = rows;
    return BI_OK;
}

// Chooses the multiply's convolution: 27-bit limbs (m <= 864), 32-bit digits,
// or (opt-in, never chosen by the tuner) bytes on the matrix cores
// (33 <= m <= 1024).
int bi_tune_set_mul_limb_bits(bi_ctx* ctx, uint32_t bits) {
    if (!ctx || (bits != 8 && bits != 27 && bits != 32)) return BI_ERR_ARG;
    ctx->mul_cfg.limb_bits = bits;
    return BI_OK;
}

int bi_tune_get_mul_limb_bits(bi_ctx* ctx, uint32_t* bits) {
    if (!ctx || !bits) return BI_ERR_ARG;
    *bits = ctx->mul_cfg.limb_bits;
    return BI_OK;
}

// Text table: "add <class> <span> <tiles> <nt>", "sub <class> <span> <tiles> <nt>",
// "mul <wpb> <rows>", "mullimb <bits>", "csr <lanes>", "csrrows <rows>",
// "csradd <span> <tiles> <nt>".
int bi_tune_save(bi_ctx* ctx, const char* path) {
    if (!ctx || !path) return BI_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) return BI_ERR_ARG;
    fprintf(f, "# libbigint tuned shapes\n");
    for (int c = 0; c < 16; c++)
        fprintf(f, "add %d %u %u %u\n", c, ctx->add_cfg[c].span_digits, ctx->add_cfg[c].tiles_in_flight,
                ctx->add_cfg[c].nontemporal);
    for (int c = 0; c < 16; c++)
        fprintf(f, "sub %d %u %u %u\n", c, ctx->sub_cfg[c].span_digits, ctx->sub_cfg[c].tiles_in_flight,
                ctx->sub_cfg[c].nontemporal);
    fprintf(f, "mul %u %u\n", ctx->mul_cfg.waves_per_block, ctx->mul_cfg.rows);
    fprintf(f, "mullimb %u\n", ctx->mul_cfg.limb_bits);
    fprintf(f, "csr %u\n", ctx->csr_lanes);
    fprintf(f, "csrrows %u\n", ctx->csr_rows);
    fprintf(f, "csradd %u %u %u\n", ctx->csr_add_cfg.span_digits, ctx->csr_add_cfg.tiles_in_flight,
            ctx->csr_add_cfg.nontemporal);
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
        if (sscanf(line, "add %d %u %u %u", &c, &a, &b, &d) == 4 && c >= 0 && c < 16 && add_shape_ok(a, b)) {
            ctx->add_cfg[c] = add_shape(a, b, d);
        } else if (sscanf(line, "sub %d %u %u %u", &c, &a, &b, &d) == 4 && c >= 0 && c < 16 && add_shape_ok(a, b)) {
            ctx->sub_cfg[c] = add_shape(a, b, d);
        } else if (sscanf(line, "mul %u %u", &a, &b) == 2 && a >= 1 && a <= 16 && (b == 4 || b == 8)) {
            ctx->mul_cfg.waves_per_block = a;
            ctx->mul_cfg.rows = b;
        } else if (sscanf(line, "mullimb %u", &a) == 1 && (a == 8 || a == 27 || a == 32)) {
            ctx->mul_cfg.limb_bits = a;
        } else if (sscanf(line, "csrrows %u", &a) == 1 && (a == 4 || a == 8)) {
            ctx->csr_rows = a;
        } else if (sscanf(line, "csradd %u %u %u", &a, &b, &d) == 3 && add_shape_ok(a, b)) {
            ctx->csr_add_cfg = add_shape(a, b, d);
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

int bi_tune_get_sub(bi_ctx* ctx, uint32_t m, uint32_t* span_digits, uint32_t* tiles_in_flight,
                    uint32_t* nontemporal) {
    if (!ctx) return BI_ERR_ARG;
    const AddCfg& c = ctx->sub_cfg_for(m ? m : 1);
    if (span_digits) *span_digits = c.span_digits;
    if (tiles_in_flight) *tiles_in_flight = c.tiles_in_flight;
    if (nontemporal) *nontemporal = c.nontemporal;
    return BI_OK;
}

int bi_tune_get_csr(bi_ctx* ctx, uint32_t* add_span_digits, uint32_t* add_tiles_in_flight, uint32_t* add_nontemporal,
                    uint32_t* mul_lanes, uint32_t* mul_rows) {
    if (!ctx) return BI_ERR_ARG;
    if (add_span_digits) *add_span_digits = ctx->csr_add_cfg.span_digits;
    if (add_tiles_in_flight) *add_tiles_in_flight = ctx->csr_add_cfg.tiles_in_flight;
    if (add_nontemporal) *add_nontemporal = ctx->csr_add_cfg.nontemporal;
    if (mul_lanes) *mul_lanes = ctx->csr_lanes;
    if (mul_rows) *mul_rows = ctx->csr_rows;
    return BI_OK;
}

}  // extern "C"
