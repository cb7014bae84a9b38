/*
 * bigint_oracle.h -- sequential CPU restatement of the reference's big-integer
 * algorithms, used ONLY as a checker (tests, smoke check, CPU baseline leg of
 * the benchmark).  Nothing in the product library links or calls this code.
 *
 * Representation (reference: thesis/sections/bigint.tex:8-26, cuda/main.cu:119):
 *   a big integer of m digits is an array of m uint32_t, little-endian
 *   (u = sum u[i] * B^i, B = 2^32); a batch is [n][m], instance-major.
 *   u64-limb batches are the same bytes viewed with m/2 limbs
 *   (cuda/main.cu:255-258, cuda/helpers.h:152-158).
 *
 * Every routine is a plain schoolbook restatement; see the per-function
 * comments for the reference lines each one follows.
 */
#ifndef BIGINT_ORACLE_H
#define BIGINT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- single-instance primitives (m digits each) ------------------------ */
/* (u+v) mod B^m, returns the carry out.  cuda/ker-add.cu.h:57-66 */
uint32_t or_add(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w);
/* |u-v| and sign (1 iff u<v).  futhark/sub.fut:14-26 */
int or_sub(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w);
/* low m digits of u*v.  cuda/ker-mul.cu.h:160-185, cuda/helpers.h:161-187 */
void or_mul_low(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w);
/* full product: w has nu+nv digits.  (GMP mpz_mul semantics) */
void or_mul_full(uint32_t nu, const uint32_t *u, uint32_t nv, const uint32_t *v, uint32_t *w);
/* comparisons.  prototypes/div.c:49-69, futhark/helper.fut:83-91 */
int or_lt(uint32_t m, const uint32_t *u, const uint32_t *v);
int or_eq(uint32_t m, const uint32_t *u, const uint32_t *v);
int or_ez(uint32_t m, const uint32_t *u);
/* -1 / 0 / +1 */
int or_cmp(uint32_t m, const uint32_t *u, const uint32_t *v);
/* index of the most significant nonzero digit (0 for zero).  div.c:92-96 */
uint32_t or_findk(uint32_t m, const uint32_t *u);
/* findk+1.  div.c:99-101 */
uint32_t or_prec(uint32_t m, const uint32_t *u);
/* v = u * B^s (s>=0, toward more significant) or floor(u / B^-s) (s<0),
 * truncated to m digits.  div.c:105-117, futhark/helper.fut:96-98 */
void or_shift(int32_t s, uint32_t m, const uint32_t *u, uint32_t *v);
/* (u*d) mod B^m.  div.c:202-218 */
void or_muld(uint32_t m, const uint32_t *u, uint32_t d, uint32_t *w);
/* floor(u/d), returns u mod d; d != 0.  div.c:165-181 */
uint32_t or_quod(uint32_t m, const uint32_t *u, uint32_t d, uint32_t *w);
/* whole-shifted inverse shinv_h(v) = floor(B^h / v) within 1 (Watt), on a
 * p-digit working precision; div.c:352-423.  Returns 0, or -1 if v == 0. */
int or_shinv(const uint32_t *v, int h, int k, uint32_t *w, uint32_t p);
/* q = floor(u/v), r = u mod v (m digits each).  div.c:427-467.
 * Returns 0, or -1 if v == 0 (q, r then zeroed). */
int or_div(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *q, uint32_t *r);

/* ---- batched drivers over [n][m] (OpenMP over instances) --------------- */
void or_add_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs, int threads);
void or_sub_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs, uint8_t *sign, int threads);
void or_mul_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs, int threads);
void or_cmp_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, int8_t *out, int threads);
void or_prec_batch(uint32_t m, uint64_t n, const uint32_t *as, uint32_t *out, int threads);
void or_shift_batch(int32_t s, uint32_t m, uint64_t n, const uint32_t *as, uint32_t *rs, int threads);
void or_muld_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *ds, uint32_t *rs, int threads);
void or_quod_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *ds, uint32_t *rs, uint32_t *rem, int threads);
int  or_div_batch(uint32_t mu, uint32_t mv, uint64_t n, const uint32_t *us, const uint32_t *vs,
                  uint32_t *qs, uint32_t *rs, int threads);

/* ---- ragged CSR batches: instance i occupies digits [off[i], off[i+1]) -- */
void or_add_csr(uint64_t n, const int64_t *off, const uint32_t *as, const uint32_t *bs, uint32_t *rs, int threads);
/* mode 0: low n_i digits of a_i*b_i into [off[i],off[i+1]);
 * mode 1: full 2 n_i digit product into [off_out[i], off_out[i+1]) */
void or_mul_csr(uint64_t n, const int64_t *off, const int64_t *off_out, const uint32_t *as,
                const uint32_t *bs, uint32_t *rs, int mode, int threads);

/* ---- GNU MP over whole batches, OpenMP over instances (gmp_batch.c) ----
 * The reference's CPU path (cuda/helpers.h:161-187, cuda/main.cu:225-245,
 * 422-442; prototypes/div.c:470-486).  op 0 add, 1 multiply (both mod
 * B^m_out), 2 divide (q: m_out = ma digits into rs, r: mb digits into rs2).
 * Return 0, -1 without libgmp.so.10, -2 after a zero divisor. */
int og_available(void);
int og_batch(int op, uint32_t ma, uint32_t mb, uint32_t m_out, uint64_t n, const uint32_t *as, const uint32_t *bs,
             uint32_t *rs, uint32_t *rs2, int threads);
/* ragged: op 0 add, 1 low-n_i multiply of instance i in [off[i], off[i+1]) */
int og_batch_csr(int op, uint64_t n, const int64_t *off, const uint32_t *as, const uint32_t *bs, uint32_t *rs,
                 int threads);

/* ---- the 2-bit carry operator of cuda/ker-add.cu.h:10-23 --------------- */
uint32_t or_carry_op(uint32_t c1, uint32_t c2);

#ifdef __cplusplus
}
#endif
#endif
