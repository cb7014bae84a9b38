/*
 * bigint_oracle.c -- ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * A sequential C restatement of the reference's big-integer algorithms, used
 * by tests/, the smoke check and the CPU-baseline leg of bench.py as the
 * checker.  The product library never links, loads or calls this file.
 *
 * Pinning (how each function is checked, beyond restating the reference):
 *   or_add / or_sub / or_mul_* / or_muld / or_quod / or_shift / or_cmp /
 *   or_prec   -- closed forms via Python int and GMP (mpz_add, mpz_sub+abs,
 *                mpz_mul, mpz_fdiv_q) on seeded random and edge inputs
 *                (all-ones chains, zeros, odd limbs), tests/test_oracle.py.
 *   or_carry_op -- exhaustive associativity + neutral-element table
 *                (thesis/sections/appendix.tex:52-62), tests/test_oracle.py.
 *   or_shinv  -- known-answer vector of thesis/sections/division.tex:110-131
 *                plus the closed form floor(B^h/v) - {0,1}.
 *   or_div    -- the reference's known-answer vectors (prototypes/Makefile:10,
 *                futhark/div.fut:185-186) and GMP mpz_tdiv_qr / Python divmod.
 */
#include "bigint_oracle.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* carry operator: bit0 = carry out, bit1 = digit is all-ones (propagates).  */
/* cuda/ker-add.cu.h:10-23 (identity 2).                                     */
uint32_t or_carry_op(uint32_t c1, uint32_t c2) {
    return (c1 & c2 & 2u) | (((c1 & (c2 >> 1)) | c2) & 1u);
}

/* ------------------------------------------------------------------------ */
/* (u+v) mod B^m with a sequential carry chain.  Semantics of baddKer*      */
/* (cuda/ker-add.cu.h:57-66): the carry out of the top digit is dropped.     */
uint32_t or_add(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w) {
    uint64_t c = 0;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t s = (uint64_t)u[i] + v[i] + c;
        w[i] = (uint32_t)s;
        c = s >> 32;
    }
    return (uint32_t)c;
}

static uint32_t o_subraw(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w) {
    uint32_t br = 0;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t d = (uint64_t)u[i] - v[i] - br;
        w[i] = (uint32_t)d;
        br = (uint32_t)((d >> 63) & 1u);
    }
    return br;
}

/* lt: most significant differing digit decides.  div.c:49-55 */
int or_lt(uint32_t m, const uint32_t *u, const uint32_t *v) {
    for (uint32_t i = m; i-- > 0;) {
        if (u[i] != v[i]) return u[i] < v[i];
    }
    return 0;
}
int or_eq(uint32_t m, const uint32_t *u, const uint32_t *v) {
    for (uint32_t i = 0; i < m; i++) if (u[i] != v[i]) return 0;
    return 1;
}
int or_ez(uint32_t m, const uint32_t *u) {
    for (uint32_t i = 0; i < m; i++) if (u[i]) return 0;
    return 1;
}
int or_cmp(uint32_t m, const uint32_t *u, const uint32_t *v) {
    for (uint32_t i = m; i-- > 0;) {
        if (u[i] != v[i]) return u[i] < v[i] ? -1 : 1;
    }
    return 0;
}

/* |u-v| with sign = (u < v): futhark/sub.fut:14-26 picks the larger operand
 * with lt and subtracts the smaller one. */
int or_sub(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w) {
    int sign = or_lt(m, u, v);
    if (sign) o_subraw(m, v, u, w); else o_subraw(m, u, v, w);
    return sign;
}

/* full product, product scanning with a 128-bit column accumulator */
void or_mul_full(uint32_t nu, const uint32_t *u, uint32_t nv, const uint32_t *v, uint32_t *w) {
    uint32_t nw = nu + nv;
    unsigned __int128 acc = 0;
    for (uint32_t k = 0; k < nw; k++) {
        uint32_t i0 = (k >= nv) ? k - nv + 1 : 0;
        uint32_t i1 = (k < nu) ? k : nu - 1;
        if (nu && nv && k < nu + nv - 1) {
            for (uint32_t i = i0; i <= i1; i++) acc += (uint64_t)u[i] * v[k - i];
        }
        w[k] = (uint32_t)acc;
        acc >>= 32;
    }
}

/* low m digits of u*v: the columns k < m only (cuda/ker-mul.cu.h:160-185
 * computes exactly the convolutions k = 0..m-1). */
void or_mul_low(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *w) {
    unsigned __int128 acc = 0;
    for (uint32_t k = 0; k < m; k++) {
        for (uint32_t i = 0; i <= k; i++) acc += (uint64_t)u[i] * v[k - i];
        w[k] = (uint32_t)acc;
        acc >>= 32;
    }
}

/* div.c:92-101 */
uint32_t or_findk(uint32_t m, const uint32_t *u) {
    for (uint32_t i = m; i-- > 0;) if (u[i]) return i;
    return 0;
}
uint32_t or_prec(uint32_t m, const uint32_t *u) { return or_findk(m, u) + 1; }

/* div.c:105-117: s >= 0 moves digits toward the more significant end. */
void or_shift(int32_t s, uint32_t m, const uint32_t *u, uint32_t *v) {
    if (s >= 0) {
        for (int64_t i = (int64_t)m - 1; i >= 0; i--) {
            int64_t off = i - s;
            v[i] = (off >= 0) ? u[off] : 0;
        }
    } else {
        for (int64_t i = 0; i < (int64_t)m; i++) {
            int64_t off = i - s;
            v[i] = (off < (int64_t)m) ? u[off] : 0;
        }
    }
}

/* div.c:202-218 */
void or_muld(uint32_t m, const uint32_t *u, uint32_t d, uint32_t *w) {
    uint64_t c = 0;
    for (uint32_t i = 0; i < m; i++) {
        uint64_t t = (uint64_t)u[i] * d + c;
        w[i] = (uint32_t)t;
        c = t >> 32;
    }
}

/* div.c:165-181 (long division by one digit) */
uint32_t or_quod(uint32_t m, const uint32_t *u, uint32_t d, uint32_t *w) {
    uint64_t r = 0;
    for (int64_t i = (int64_t)m - 1; i >= 0; i--) {
        r = (r << 32) | u[i];
        w[i] = (uint32_t)(r / d);
        r = r % d;
    }
    return (uint32_t)r;
}

/* ======================================================================== */
/* Division by whole-shifted inverse, restating prototypes/div.c:221-467.   */
/* All big integers live in p-digit buffers; products are computed on their  */
/* significant digits and truncated to the buffer (the reference's buffers    */
/* are sized so that no product is ever truncated).                           */
/* ======================================================================== */

static void o_zero(uint32_t *u, uint32_t m) { memset(u, 0, (size_t)m * 4); }
static void o_set(uint32_t *u, uint32_t d, uint32_t m) { o_zero(u, m); u[0] = d; }
static void o_bpow(int n, uint32_t m, uint32_t *b) {   /* div.c:121-126 */
    o_zero(b, m);
    if (n >= 0 && (uint32_t)n < m) b[n] = 1;
}
/* w = (u mod B^d) * (v mod B^d), truncated to m digits (w may alias u or v).
 * With d == m this is div.c:184-199's mult; with d < m it is the product
 * that div.c:221-242's multmod exports (the full product of the truncated
 * operands). */
static void o_mult_d(const uint32_t *u, const uint32_t *v, int d, uint32_t *w, uint32_t m) {
    if (d <= 0) { o_zero(w, m); return; }
    uint32_t nu = or_prec((uint32_t)d, u), nv = or_prec((uint32_t)d, v);
    uint32_t *t = (uint32_t *)calloc((size_t)nu + nv, 4);
    or_mul_full(nu, u, nv, v, t);
    o_zero(w, m);
    uint32_t nt = nu + nv < m ? nu + nv : m;
    memcpy(w, t, (size_t)nt * 4);
    free(t);
}
/* GMP-backed add/sub of div.c:129-162: the full (or absolute) result,
 * truncated to the buffer */
static void o_addm(const uint32_t *u, const uint32_t *v, uint32_t *w, uint32_t m) { or_add(m, u, v, w); }
static void o_absdiff(const uint32_t *u, const uint32_t *v, uint32_t *w, uint32_t m) { (void)or_sub(m, u, v, w); }

/* div.c:246-280 */
static int o_powerdiff(const uint32_t *v, const uint32_t *w, int h, int l, uint32_t *P, uint32_t m) {
    int L = (int)or_prec(m, v) + (int)or_prec(m, w) - l + 1;
    uint32_t *bh = (uint32_t *)malloc((size_t)m * 4);
    o_bpow(h, m, bh);
    int sign = 0;
    if (or_ez(m, v) || or_ez(m, w)) {
        memcpy(P, bh, (size_t)m * 4);
    } else if (L >= h) {
        o_mult_d(v, w, (int)m, P, m);
        if (or_lt(m, P, bh)) o_absdiff(bh, P, P, m);
        else { o_absdiff(P, bh, P, m); sign = 1; }
    } else {
        o_mult_d(v, w, L, P, m);
        if (!or_ez(m, P)) {
            if (P[L - 1] == 0) sign = 1;
            else {
                uint32_t *bL = (uint32_t *)malloc((size_t)m * 4);
                o_bpow(L, m, bL);
                if (or_lt(m, P, bL)) o_absdiff(bL, P, P, m);
                else { o_absdiff(P, bL, P, m); sign = 1; }
                free(bL);
            }
        }
    }
    free(bh);
    return sign;
}

/* div.c:283-304 */
static void o_step(int h, const uint32_t *v, uint32_t *w, int n, int l, int g, uint32_t m) {
    uint32_t *P = (uint32_t *)calloc(m, 4);
    uint32_t *w0 = (uint32_t *)malloc((size_t)m * 4);
    or_shift(n, m, w, w0);
    int sign = o_powerdiff(v, w, h - n, l - g, P, m);
    o_mult_d(w, P, (int)m, w, m);
    or_shift(2 * n - h, m, w, w);
    if (sign) o_absdiff(w, w0, w, m);
    else o_addm(w, w0, w, m);
    free(w0); free(P);
}

/* div.c:333-349 (refine3, g = m/4 guard digits on the 4x buffer) */
static void o_refine3(const uint32_t *v, int h, int k, uint32_t *w, int l, uint32_t m) {
    int g = (int)(m / 4);
    or_shift(g, m, w, w);
    uint32_t *v0 = (uint32_t *)calloc(m, 4);
    while (h - k > l) {
        int n = (h - k + 1 - l) < l ? (h - k + 1 - l) : l;
        int s = k - 2 * l + 1 - g; if (s < 0) s = 0;
        or_shift(-s, m, v, v0);
        o_step(k + l + n - s + g, v0, w, n, l, g, m);
        or_shift(-1, m, w, w);
        l += n - 1;
    }
    or_shift(-g, m, w, w);
    free(v0);
}

/* div.c:352-423 */
int or_shinv(const uint32_t *v, int h, int k, uint32_t *w, uint32_t m) {
    uint32_t *b = (uint32_t *)malloc((size_t)m * 4), *bh = (uint32_t *)malloc((size_t)m * 4);
    uint32_t *bk = (uint32_t *)malloc((size_t)m * 4), *v2 = (uint32_t *)malloc((size_t)m * 4);
    o_bpow(1, m, b); o_bpow(h, m, bh); o_bpow(k, m, bk);
    or_muld(m, v, 2, v2);
    int done = 1, rc = 0;
    if (or_lt(m, v, b)) {
        if (v[0] == 0) { o_zero(w, m); rc = -1; }
        else or_quod(m, bh, v[0], w);
    } else if (or_lt(m, bh, v)) {
        o_set(w, 0, m);
    } else if (or_lt(m, bh, v2)) {
        o_set(w, 1, m);
    } else if (or_eq(m, v, bk)) {
        o_set(w, 1, m);
        or_shift(h - k, m, w, w);
    } else {
        done = 0;
    }
    free(b); free(bh); free(bk); free(v2);
    if (done) return rc;

    /* initial approximation from the three leading digits (div.c:391-418):
     * W = (2^128 - V) / V + 1, V = v[k-2] + v[k-1] B + v[k] B^2 */
    int l = 2;
    unsigned __int128 V = (unsigned __int128)v[k - 2] | ((unsigned __int128)v[k - 1] << 32)
                        | ((unsigned __int128)v[k] << 64);
    unsigned __int128 r = ((unsigned __int128)0 - V) / V + 1;
    o_zero(w, m);
    w[0] = (uint32_t)r; w[1] = (uint32_t)(r >> 32);
    w[2] = (uint32_t)(r >> 64); w[3] = (uint32_t)(r >> 96);
    if (h - k <= l) or_shift(h - k - l, m, w, w);
    else o_refine3(v, h, k, w, l, m);
    return 0;
}

/* div.c:427-467 */
int or_div(uint32_t m, const uint32_t *u, const uint32_t *v, uint32_t *q, uint32_t *r) {
    if (or_ez(m, v)) { o_zero(q, m); o_zero(r, m); return -1; }
    int h = (int)or_findk(m, u) + 1;
    int k = (int)or_findk(m, v);
    uint32_t p = m * 4;
    uint32_t *a = (uint32_t *)calloc(p, 4), *b = (uint32_t *)calloc(p, 4), *c = (uint32_t *)calloc(p, 4);
    memcpy(a, u, (size_t)m * 4);
    memcpy(b, v, (size_t)m * 4);
    if (k == 1) {
        h++; k++;
        or_shift(1, m + 1, a, a);
        or_shift(1, m + 1, b, b);
    }
    or_shinv(b, h, k, c, p);
    o_mult_d(a, c, (int)p, b, p);
    or_shift(-h, p, b, b);
    memcpy(q, b, (size_t)m * 4);
    o_mult_d(v, q, (int)m, a, m);
    o_absdiff(u, a, r, m);
    if (!or_lt(m, r, v)) {
        o_set(a, 1, m);
        or_add(m, q, a, q);
        o_absdiff(r, v, r, m);
    }
    free(a); free(b); free(c);
    return 0;
}

/* ======================================================================== */
/* batched drivers                                                           */
/* ======================================================================== */
static int o_threads(int t) {
#ifdef _OPENMP
    return t > 0 ? t : omp_get_max_threads();
#else
    (void)t; return 1;
#endif
}

void or_add_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) or_add(m, as + i * m, bs + i * m, rs + i * m);
}
void or_sub_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs, uint8_t *sign, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) sign[i] = (uint8_t)or_sub(m, as + i * m, bs + i * m, rs + i * m);
}
void or_mul_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs, int threads) {
    #pragma omp parallel for schedule(dynamic, 16) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) {
        /* low m digits via the full product of the significant digits
         * (identical result to or_mul_low, far cheaper on top-zero data) */
        const uint32_t *a = as + i * m, *b = bs + i * m;
        uint32_t na = or_prec(m, a), nb = or_prec(m, b);
        uint32_t *t = (uint32_t *)malloc(((size_t)na + nb) * 4);
        or_mul_full(na, a, nb, b, t);
        uint32_t nt = na + nb < m ? na + nb : m;
        memcpy(rs + i * m, t, (size_t)nt * 4);
        memset(rs + i * m + nt, 0, (size_t)(m - nt) * 4);
        free(t);
    }
}
/* full products of n pairs (ma and mb digits -> ma + mb digits): the
 * schoolbook of or_mul_full per pair (the long products the reference
 * leaves to mpz_mul, prototypes/div.c:221-242) */
void or_mul_full_batch(uint32_t ma, uint32_t mb, uint64_t n, const uint32_t *as, const uint32_t *bs, uint32_t *rs,
                       int threads) {
    #pragma omp parallel for schedule(dynamic, 1) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++)
        or_mul_full(ma, as + i * ma, mb, bs + i * mb, rs + i * ((size_t)ma + mb));
}
void or_cmp_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *bs, int8_t *out, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) out[i] = (int8_t)or_cmp(m, as + i * m, bs + i * m);
}
void or_prec_batch(uint32_t m, uint64_t n, const uint32_t *as, uint32_t *out, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) out[i] = or_prec(m, as + i * m);
}
void or_shift_batch(int32_t s, uint32_t m, uint64_t n, const uint32_t *as, uint32_t *rs, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) or_shift(s, m, as + i * m, rs + i * m);
}
void or_muld_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *ds, uint32_t *rs, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) or_muld(m, as + i * m, ds[i], rs + i * m);
}
void or_quod_batch(uint32_t m, uint64_t n, const uint32_t *as, const uint32_t *ds, uint32_t *rs, uint32_t *rem, int threads) {
    #pragma omp parallel for schedule(static) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) rem[i] = or_quod(m, as + i * m, ds[i], rs + i * m);
}
/* u: [n][mu], v: [n][mv] (mv <= mu); q: [n][mu], r: [n][mv] */
int or_div_batch(uint32_t mu, uint32_t mv, uint64_t n, const uint32_t *us, const uint32_t *vs,
                 uint32_t *qs, uint32_t *rs, int threads) {
    int bad = 0;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(o_threads(threads)) reduction(|:bad)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t *v = (uint32_t *)calloc(mu, 4), *r = (uint32_t *)calloc(mu, 4);
        memcpy(v, vs + i * mv, (size_t)mv * 4);
        if (or_div(mu, us + i * (uint64_t)mu, v, qs + i * (uint64_t)mu, r) != 0) bad = 1;
        memcpy(rs + i * (uint64_t)mv, r, (size_t)mv * 4);
        free(v); free(r);
    }
    return bad ? -1 : 0;
}

void or_add_csr(uint64_t n, const int64_t *off, const uint32_t *as, const uint32_t *bs, uint32_t *rs, int threads) {
    #pragma omp parallel for schedule(dynamic, 64) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) {
        int64_t o = off[i];
        or_add((uint32_t)(off[i + 1] - o), as + o, bs + o, rs + o);
    }
}
void or_mul_csr(uint64_t n, const int64_t *off, const int64_t *off_out, const uint32_t *as,
                const uint32_t *bs, uint32_t *rs, int mode, int threads) {
    #pragma omp parallel for schedule(dynamic, 16) num_threads(o_threads(threads))
    for (int64_t i = 0; i < (int64_t)n; i++) {
        int64_t o = off[i];
        uint32_t m = (uint32_t)(off[i + 1] - o);
        const uint32_t *a = as + o, *b = bs + o;
        uint32_t na = or_prec(m, a), nb = or_prec(m, b);
        uint32_t *t = (uint32_t *)malloc(((size_t)na + nb) * 4);
        or_mul_full(na, a, nb, b, t);
        uint32_t *r = rs + (mode ? off_out[i] : o);
        uint32_t mr = mode ? 2 * m : m;
        uint32_t nt = na + nb < mr ? na + nb : mr;
        memcpy(r, t, (size_t)nt * 4);
        memset(r + nt, 0, (size_t)(mr - nt) * 4);
        free(t);
    }
}
