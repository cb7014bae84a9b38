/*
 * bigint.h -- C ABI of libbigint, a batched exact big-integer engine for
 * AMD Instinct MI355X (gfx950).
 *
 * Integers are arrays of little-endian uint32_t digits (B = 2^32); a batch is
 * [n][m] instance-major; a u64-limb batch is the same bytes with m/2 limbs.
 * This is the layout of the reference's batched entry points
 * (cuda/main.cu:108-110,280-282; cuda/helpers.h:152-158).
 *
 * Conventions (replacing the reference's void/assert/printf behaviour,
 * cuda/main.cu:112-113,131-135,162,199):
 *   - every call returns BI_OK (0) or a negative BI_ERR_* code; nothing exits;
 *   - any n is accepted (no tail instance is dropped, unlike cuda/main.cu:135);
 *   - the caller owns every buffer; a context owns one HIP stream (default
 *     priority), a grow-only device scratch arena and its tuned kernel shapes;
 *   - *_dev entry points take device pointers and are asynchronous on the
 *     context stream; the others take host pointers and return when the
 *     results are in host memory;
 *   - outputs may alias inputs of the same shape (each kernel reads a digit
 *     before any lane writes it);
 *   - one context per host thread; distinct contexts may run concurrently.
 */
#ifndef BIGINT_H
#define BIGINT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BI_OK 0
#define BI_ERR_ARG (-1)      /* null pointer, zero size, bad mode */
#define BI_ERR_SHAPE (-2)    /* size outside what the kernels support */
#define BI_ERR_HIP (-3)      /* HIP runtime error (see bi_last_hip_error) */
#define BI_ERR_NOMEM (-4)    /* device allocation failed / budget exceeded */
#define BI_ERR_DIVZERO (-5)  /* a divisor is zero */
#define BI_ERR_BUSY (-6)     /* serving: no free request slot */
#define BI_ERR_NODEV (-7)    /* no usable GPU */
#define BI_ERR_STATE (-8)    /* call not valid in the current state */

typedef struct bi_ctx bi_ctx;

/* ---- context (no counterpart in the reference: it allocates per call,
 *      cuda/main.cu:121-124,218-221) ---------------------------------------- */
int bi_create(bi_ctx **ctx, int device);
int bi_destroy(bi_ctx *ctx);
const char *bi_strerror(int code);
const char *bi_last_hip_error(bi_ctx *ctx);
int bi_sync(bi_ctx *ctx);
void *bi_stream(bi_ctx *ctx); /* the context's hipStream_t */

/* ---- device memory and timing plumbing (so callers need no HIP headers) -- */
int bi_dev_alloc(bi_ctx *ctx, void **ptr, size_t bytes);
int bi_dev_free(bi_ctx *ctx, void *ptr);
int bi_copy_h2d(bi_ctx *ctx, void *dst, const void *src, size_t bytes);  /* synchronous */
int bi_copy_d2h(bi_ctx *ctx, void *dst, const void *src, size_t bytes);  /* synchronous */
int bi_copy_d2d(bi_ctx *ctx, void *dst, const void *src, size_t bytes);  /* async on stream */
int bi_timer_start(bi_ctx *ctx);                 /* event recorded on the context stream */
int bi_timer_stop(bi_ctx *ctx, float *ms);       /* second event; waits; elapsed ms */

/* ---- measurement helpers (no counterpart in the reference) -------------- */
int bi_device_info(bi_ctx *ctx, int *cus, int *clock_khz, char *name, size_t name_len);
/* issue rate of the multiply's v_mad_u64_u32 + carry pair, all CUs busy */
int bi_probe_mac_rate(bi_ctx *ctx, double *macs_per_s);
/* plain nontemporal copy of `bytes` (multiple of 16, 16-byte aligned),
 * async on the context stream: the HBM calibration kernel */
int bi_copy_kernel(bi_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- kernel shapes and the autotuner (no counterpart in the reference,
 * whose shapes are compile-time rules, cuda/main.cu:131-135,303-307).
 * Shapes never change results.  what: 0 = add stream (span digits, tiles in
 * flight, nontemporal) for the size class of m; 1 = uniform multiply
 * workgroup size, rows per block step (4 or 8) and limb width; 2 = ragged
 * add (span, tiles) and ragged multiply (lanes per bundle, rows) on a sample
 * of digit counts log-uniform in [8, m] (m <= 4096); 3 = |a-b| stream for
 * the size class of m.  Candidates are timed in interleaved rounds and
 * scored by their median; the shape in place is kept unless one is 2 %
 * faster.  Table file: "add <class> <span> <tiles> <nt>", "sub <class> <span>
 * <tiles> <nt>", "mul <waves> <rows>", "mullimb <bits>", "csr <lanes>",
 * "csrrows <rows>", "csradd <span> <tiles> <nt>". */
int bi_tune_run(bi_ctx *ctx, uint32_t m, uint64_t bytes_per_operand, int what, float *best_ms);
int bi_tune_set_add(bi_ctx *ctx, uint32_t m, uint32_t span_digits, uint32_t tiles_in_flight, uint32_t nontemporal);
int bi_tune_set_sub(bi_ctx *ctx, uint32_t m, uint32_t span_digits, uint32_t tiles_in_flight, uint32_t nontemporal);
int bi_tune_set_mul(bi_ctx *ctx, uint32_t waves_per_block, uint32_t rows);
int bi_tune_set_csr_lanes(bi_ctx *ctx, uint32_t lanes);
/* tiles in flight: 1, 2, 4 or 8 (8 counts as 4 for the uniform add and |a-b|);
 * the ragged add's span is its fixed chunk of the flat digit array (rounded
 * down to whole 256-digit tiles, at least one) */
int bi_tune_set_csr(bi_ctx *ctx, uint32_t add_span_digits, uint32_t add_tiles_in_flight, uint32_t add_nontemporal,
                    uint32_t mul_lanes, uint32_t mul_rows);
/* uniform multiply convolution: 27 = 27-bit limbs (m <= 864 digits: one
 * v_mad_u64_u32 per limb product, no overflow counts), 32 = 32-bit digits
 * (any m; one mad plus one carry count per digit product), 8 = bytes on the
 * int8 matrix cores (33 <= m <= 1024 digits: operands rewritten as balanced
 * signed bytes, column sums by v_mfma_i32_32x32x32_i8) -- an opt-in beside
 * the multiply-add kernels, never picked by the tuner.  Outside a choice's
 * range the 32-bit digits are used.  Default 32; bi_tune_run(what = 1)
 * times 27 and 32 where both apply.  Table line "mullimb <bits>". */
int bi_tune_set_mul_limb_bits(bi_ctx *ctx, uint32_t bits);
int bi_tune_get_mul_limb_bits(bi_ctx *ctx, uint32_t *bits);
int bi_tune_get(bi_ctx *ctx, uint32_t m, uint32_t *span_digits, uint32_t *tiles_in_flight, uint32_t *nontemporal,
                uint32_t *mul_waves, uint32_t *mul_rows, uint32_t *csr_lanes);
int bi_tune_get_sub(bi_ctx *ctx, uint32_t m, uint32_t *span_digits, uint32_t *tiles_in_flight,
                    uint32_t *nontemporal);
int bi_tune_get_csr(bi_ctx *ctx, uint32_t *add_span_digits, uint32_t *add_tiles_in_flight, uint32_t *add_nontemporal,
                    uint32_t *mul_lanes, uint32_t *mul_rows);
int bi_tune_save(bi_ctx *ctx, const char *path);
int bi_tune_load(bi_ctx *ctx, const char *path);

/* ---- plain batched add: rs = (as + bs) mod B^m ----------------------------
 * replaces gpuAdd<U32bits|U64bits, m, v>(n, h_as, h_bs, h_rs)
 * (cuda/main.cu:108-222; kernels cuda/ker-add.cu.h:50-266). */
int bi_add_u32(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs);
int bi_add_u64(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, const uint64_t *bs, uint64_t *rs);
int bi_add_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs);
int bi_add_u64_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, const uint64_t *bs, uint64_t *rs);

/* ---- plain batched multiply: rs = low m limbs of as*bs --------------------
 * replaces gpuMultiply<U32bits|U64bits, m, v>(n, h_as, h_bs, h_rs)
 * (cuda/main.cu:280-419; kernels cuda/ker-mul.cu.h:147-666).
 * m <= 4096 u32 digits (2048 u64 limbs), the reference's own limit
 * (m/q <= 1024 threads with q = 4, cuda/main.cu:303-307). */
int bi_mul_u32(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs);
int bi_mul_u64(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, const uint64_t *bs, uint64_t *rs);
int bi_mul_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs);
int bi_mul_u64_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, const uint64_t *bs, uint64_t *rs);
/* The same multiply as bi_mul_u32_dev (same kernels, same results), calling
 * no allocator: no hipMalloc, hipFree or allocating library routine.  No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its scratch is a static device array of 24 MiB per GPU
 * (the multiply's per-lane snapshots for 2048 lanes on each of 256 CUs; its
 * size does not grow with n or m), held until the process ends and shared by
 * every call on that GPU: calls that overlap in time on one GPU race and must
 * be ordered by the caller.  m <= 4096 (BI_ERR_SHAPE above);
 * BI_ERR_STATE if the array's device address could not be resolved when the
 * context was created. */
int bi_mul_u32_dev_noalloc(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs);

/* ---- full products of long operands (device pointers) --------------------
 * rs = as * bs, all ma + mb digits, for n pairs (as: [n][ma], bs: [n][mb],
 * rs: [n][ma+mb]).  The reference multiplies long operands only through GMP:
 * prototypes/div.c:184-199 (`mult`, a stand-in for a convolution or transform
 * product, as its comment says) and the
 * oracle cuda/helpers.h:161-180 (`gmpAddMulOnce`); this entry point is the
 * batched replacement, with two algorithms:
 *   algo 1: quadratic long product (VALU multiply-adds), ma, mb up to about
 *           20000 digits together;
 *   algo 2: number-theoretic transform over p = 2^64 - 2^32 + 1 (16- or
 *           25-bit chunks, one workgroup per product), up to 2^14 points:
 *           ma + mb <= 8192, or up to 12800 when 25-bit chunks fit
 *           (ceil(32 ma / 25) + ceil(32 mb / 25) <= 16384);
 *   algo 0: the faster of the two by the measured crossover: the transform
 *           when ma * mb reaches 1.6e6 / 3.7e6 / 7.5e6 digit products for a
 *           2^12 / 2^13 / 2^14-point transform (e.g. from 1536 x 1536
 *           digits up), else the quadratic product.
 * BI_ERR_SHAPE when the chosen algorithm cannot hold the sizes. */
int bi_mul_full_u32_dev(bi_ctx *ctx, uint64_t n, uint32_t ma, uint32_t mb, const uint32_t *as, const uint32_t *bs,
                        uint32_t *rs, int algo);

/* ---- fused chains (device pointers) --------------------------------------
 * Replace the reference's in-kernel benchmark chains: baddKer1/2/3Bench
 * (cuda/ker-add.cu.h:88-112,175-203,268-295): rs = p*(as+bs) mod B^m as p
 * chained additions (registers for m <= 2048); convMul1..5Bench
 * (cuda/ker-mul.cu.h:223-250 and twins): rs = (as*bs)^p mod B^m as p chained
 * products (LDS for m <= 512, m <= 4096).  p >= 1. */
int bi_add_chain_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, uint32_t p, const uint32_t *as, const uint32_t *bs,
                         uint32_t *rs);
int bi_mul_chain_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, uint32_t p, const uint32_t *as, const uint32_t *bs,
                         uint32_t *rs);

/* ---- GMP-compatible export / import (device pointers) ---------------------
 * The mpz_export / mpz_import format with order -1 (least significant word
 * first), native endian, no nails, words of wsize = 4 or 8 bytes: the format
 * the reference hands integers to GMP in (cuda/helpers.h:113,167-177).
 * bi_export_sizes_u32_dev: offs[0] = 0, offs[i+1] = offs[i] + the number of
 * words mpz_export writes for instance i (its significant words, 0 for zero);
 * bi_export_u32_dev: those words, back to back at offs[i] words into out;
 * bi_import_u32_dev: the inverse into a uniform [n][m] batch (zero-padded,
 * truncated mod B^m).  offs: n+1 uint64 entries in device memory. */
int bi_export_sizes_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *xs, uint32_t wsize, uint64_t *offs);
int bi_export_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *xs, uint32_t wsize, const uint64_t *offs,
                      void *out);
int bi_import_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const void *in, uint32_t wsize, const uint64_t *offs,
                      uint32_t *xs);

/* ---- ragged (CSR) batches ------------------------------------------------
 * Instance i occupies digits [offs[i], offs[i+1]) of as, bs and rs (int64
 * offsets, n+1 entries, device memory).  No counterpart in the reference,
 * whose batches have one size per launch (cuda/main.cu:131-135,303-307). */
typedef struct bi_plan bi_plan;
/* rs = (as + bs) mod B^{n_i} per instance; total_digits = offs[n] */
int bi_add_csr_u32_dev(bi_ctx *ctx, uint64_t n, const int64_t *offs, uint64_t total_digits, const uint32_t *as,
                       const uint32_t *bs, uint32_t *rs);
/* Work list of a ragged multiply, built on the host from a HOST copy of the
 * offsets.  mode 0: low n_i digits (n_i <= 4096); mode 1: full 2 n_i digit
 * product at offs_out (n_i <= 2048).  lane_budget: lanes per workgroup
 * (multiple of 64, 64..1024; 0 = the context's tuned value). */
int bi_csr_plan_create(bi_ctx *ctx, uint64_t n, const int64_t *offs_host, int mode, uint32_t lane_budget,
                       bi_plan **plan);
int bi_csr_plan_destroy(bi_plan *plan);
int bi_csr_plan_info(const bi_plan *plan, uint32_t *nbundles, uint32_t *block, uint32_t *lds_bytes,
                     double *useful_macs);
/* ragged multiply, one persistent launch */
int bi_mul_csr_u32_dev(bi_ctx *ctx, const bi_plan *plan, const int64_t *offs, const int64_t *offs_out,
                       const uint32_t *as, const uint32_t *bs, uint32_t *rs);
/* ragged add AND multiply of the same operands in ONE persistent launch:
 * multiply bundles first (largest first), then add spans, from one counter */
int bi_addmul_csr_u32_dev(bi_ctx *ctx, const bi_plan *plan, const int64_t *offs, const int64_t *offs_out,
                          const uint32_t *as, const uint32_t *bs, uint32_t *rs_add, uint32_t *rs_mul);
/* Host-pointer forms (offsets, operands and results in host memory; one
 * round trip a call, as the reference's host entry points gpuAdd /
 * gpuMultiply, cuda/main.cu:108-128,214-221): the ragged add (offs[0] >= 0,
 * non-decreasing; digits before offs[0] keep the caller's values) and the
 * ragged multiply (builds and frees its plan; mode 0: low n_i digits at
 * offs; mode 1: 2 n_i digits at offs_out, whose windows must hold them;
 * digits outside the windows keep the caller's values).  BI_ERR_ARG for bad
 * offsets. */
int bi_add_csr_u32(bi_ctx *ctx, uint64_t n, const int64_t *offs, const uint32_t *as, const uint32_t *bs,
                   uint32_t *rs);
int bi_mul_csr_u32(bi_ctx *ctx, uint64_t n, const int64_t *offs, const int64_t *offs_out, const uint32_t *as,
                   const uint32_t *bs, uint32_t *rs, int mode);

/* ---- division -------------------------------------------------------------
 * qs = floor(us / vs) ([n][mu] digits), rs = us mod vs ([n][mv] digits), by
 * Newton iteration for the whole-shifted inverse (prototypes/div.c:352-467,
 * futhark/div.fut:134-167, thesis/sections/division.tex:140-405), captured in
 * one hipGraph per batch shape.  Synchronous: returns when q, r are written.
 * Returns BI_ERR_DIVZERO if any divisor is zero (those instances get q = r = 0,
 * all others are complete).  mu + max(mv, 3) <= 16384 digits. */
int bi_divmod_u32_dev(bi_ctx *ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t *us, const uint32_t *vs,
                      uint32_t *qs, uint32_t *rs);
int bi_divmod_u64_dev(bi_ctx *ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint64_t *us, const uint64_t *vs,
                      uint64_t *qs, uint64_t *rs);
/* The same with us, vs, qs, rs in host memory (one round trip a call; the
 * reference's div.c divides arrays in host memory, prototypes/div.c:427-467;
 * u64: mu, mv in u64 limbs, cuda/main.cu:255-258).  BI_ERR_DIVZERO as above,
 * with q and r of every other instance copied out. */
int bi_divmod_u32(bi_ctx *ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t *us, const uint32_t *vs,
                  uint32_t *qs, uint32_t *rs);
int bi_divmod_u64(bi_ctx *ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint64_t *us, const uint64_t *vs,
                  uint64_t *qs, uint64_t *rs);
/* Batches of at most n_max divisions (default 4096; at most 2048 when an
 * operand passes 2048 digits) whose shape fits one workgroup's LDS (mu up to
 * about 6700 digits) run as ONE launch: a workgroup of 256 or 1024 lanes per
 * instance takes the whole Newton pipeline above with every working buffer
 * in LDS and its own precision schedule (no counterpart in the reference,
 * whose div.c divides one instance at a time on the CPU); larger batches
 * take the batched pipeline.  n_max = 0: always the batched pipeline. */
int bi_tune_set_div_one(bi_ctx *ctx, uint32_t n_max);
int bi_tune_get_div_one(bi_ctx *ctx, uint32_t *n_max);
/* precision (quotient digits + 1), Newton steps and digit products of the
 * last division on this context (zero after a one-launch division: its
 * instances each have their own schedule) */
int bi_div_last_info(bi_ctx *ctx, uint32_t *P, uint32_t *newton_steps, double *macs);
/* the same digit products split by how they were computed: valu_macs by
 * the quadratic multiply-add kernels (v_mad_u64_u32), transform_macs the
 * quadratic equivalent of the products taken by number-theoretic transform
 * (valu_macs + transform_macs = macs above) */
int bi_div_last_work(bi_ctx *ctx, double *valu_macs, double *transform_macs);

/* ---- serving: continuous batching on one GPU (no counterpart in the
 * reference).  Requests are copied in at submission; a scheduler thread
 * admits them in arrival order while their operands and results fit the free
 * part of an HBM arena of arena_bytes, packs the admitted add/mul requests
 * into one ragged batch (one persistent launch) per free stream slot, and
 * runs divisions through the Newton pipeline; batches on different slots
 * overlap.  add/mul: a and b have ma == mb digits, result r0 = low ma digits
 * of a+b / a*b.  div: r0 = floor(a/b) (ma digits), r1 = a mod b (mb digits).
 * ma, mb <= 4096.  Divisions run on two worker streams (operands <= 512
 * digits, and longer); keep slots + 2 within the process's hardware queues
 * (GPU_MAX_HW_QUEUES, 4 by default: slots <= 2) so that no batch stream
 * shares a queue with a division stream. */
#define BI_OP_ADD 0
#define BI_OP_MUL 1
#define BI_OP_DIV 2
typedef struct bi_srv bi_srv;
int bi_srv_start(bi_ctx *ctx, size_t arena_bytes, uint32_t max_batch, uint32_t slots, bi_srv **srv);
int bi_srv_submit(bi_srv *srv, int op, uint32_t ma, const uint32_t *a, uint32_t mb, const uint32_t *b,
                  uint64_t *ticket);
/* *done = 1 once the result is ready; then copies it out, returns the
 * request's status and releases the ticket */
int bi_srv_poll(bi_srv *srv, uint64_t ticket, uint32_t *r0, uint32_t *r1, int *done, double *latency_s);
int bi_srv_stats(bi_srv *srv, uint64_t *batches, uint64_t *max_batch, size_t *peak_arena_bytes);
int bi_srv_stop(bi_srv *srv);
/* measurement driver: open-loop submission of n requests at arrival_s[i]
 * seconds, results and per-request latencies collected (see bi_serve.hip) */
int bi_srv_run_trace(bi_srv *srv, uint64_t n, const int32_t *ops, const uint32_t *ma, const uint32_t *mb,
                     const uint64_t *a_off, const uint32_t *a_pool, const uint64_t *b_off, const uint32_t *b_pool,
                     const double *arrival_s, uint32_t *r0_pool, const uint64_t *r_off0, uint32_t *r1_pool,
                     const uint64_t *r_off1, double *latency_s, int32_t *status, double *wall_s);

/* ---- helpers on the same scans (device pointers, [n][m] batches) ---------
 * These have no batched counterpart in the reference's CUDA code; they
 * restate the sequential C helpers of prototypes/div.c and the Futhark
 * specifications, batched. */
/* rs = |as - bs|, sign[i] = (as[i] < bs[i]).
 * futhark/sub.fut:14-26 (bsub); prototypes/div.c:147-162 */
int bi_sub_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs,
                   uint8_t *sign);
/* out[i] = -1, 0, 1 as as[i] <, ==, > bs[i].
 * futhark/helper.fut:83-91 (lt, eq); prototypes/div.c:49-62 */
int bi_cmp_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, int8_t *out);
/* out[i] = findk(as[i]) + 1 (1 for zero).  prototypes/div.c:92-101 */
int bi_prec_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, uint32_t *out);
/* rs = as * B^s truncated to m digits (s >= 0), or floor(as / B^-s) (s < 0).
 * prototypes/div.c:105-117, futhark/helper.fut:96-98.  rs may alias as. */
int bi_shift_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, int32_t s, const uint32_t *as, uint32_t *rs);
/* rs = (as * ds[i]) mod B^m.  prototypes/div.c:202-218, futhark/div.fut:58-61 */
int bi_muld_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *ds, uint32_t *rs);
/* rs = floor(as / ds[i]), rem[i] = as mod ds[i] (rem may be NULL).
 * prototypes/div.c:165-181.  ds[i] == 0 yields quotient 0, remainder 0. */
int bi_quod_u32_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *ds, uint32_t *rs,
                    uint32_t *rem);
/* The same helpers over u64 limbs (m limbs = the same bytes as 2m u32
 * digits; the reference's U64bits view, cuda/ker-helpers.cu.h:28-44,
 * cuda/main.cu:255-258): |a-b| with sign, compare, shift by s u64 limbs, and
 * precision counted in u64 limbs (findk over limbs + 1). */
int bi_sub_u64_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, const uint64_t *bs, uint64_t *rs,
                   uint8_t *sign);
int bi_cmp_u64_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, const uint64_t *bs, int8_t *out);
int bi_prec_u64_dev(bi_ctx *ctx, uint32_t m, uint64_t n, const uint64_t *as, uint32_t *out);
int bi_shift_u64_dev(bi_ctx *ctx, uint32_t m, uint64_t n, int32_t s, const uint64_t *as, uint64_t *rs);

#ifdef __cplusplus
}
#endif
#endif /* BIGINT_H */
