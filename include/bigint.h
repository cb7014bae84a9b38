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

#ifdef __cplusplus
}
#endif
#endif /* BIGINT_H */
