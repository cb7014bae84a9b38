/*
 * gmp_batch.c -- GNU MP over whole batches, OpenMP over instances: the CPU
 * path the reference validates and compares against, run on all host cores.
 * Test and baseline infrastructure only (part of liboracle.so); nothing in
 * the product library links or calls it.
 *
 * Per instance, as cuda/helpers.h:161-187 (gmpAddMulOnce): mpz_import of the
 * operands (order -1, 4-byte words), the GMP operation, mpz_export, result
 * truncated to the requested digit count and zero-padded.  gmpAdd /
 * gmpMultiply (cuda/main.cu:225-245,422-442) loop that over the batch on one
 * thread; here each thread keeps its mpz variables across instances (the
 * reference initialises and clears them per instance), so the figure is the
 * faster of the two.  Division is mpz_tdiv_qr, the check of
 * prototypes/div.c:470-486 (div_gmp).
 *
 * libgmp.so.10 is loaded with dlopen and called through its documented
 * entry points (__gmpz_*); the development headers are not needed.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "bigint_oracle.h"

typedef struct {
    int alloc;
    int size;
    void *d;
} og_mpz; /* mpz_t on LP64 */

typedef void (*fz)(og_mpz *);
typedef void (*fz3)(og_mpz *, const og_mpz *, const og_mpz *);
typedef void (*fqr)(og_mpz *, og_mpz *, const og_mpz *, const og_mpz *);
typedef void (*fimp)(og_mpz *, size_t, int, size_t, int, size_t, const void *);
typedef void *(*fexp)(void *, size_t *, int, size_t, int, size_t, const og_mpz *);

static struct {
    int state; /* 0 not tried, 1 loaded, -1 unavailable */
    fz init, clear;
    fz3 add, mul;
    fqr tdiv_qr;
    fimp import;
    fexp export_;
} G;

static int og_load(void) {
    if (G.state) return G.state > 0;
    void *h = dlopen("libgmp.so.10", RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        G.state = -1;
        return 0;
    }
    G.init = (fz)dlsym(h, "__gmpz_init");
    G.clear = (fz)dlsym(h, "__gmpz_clear");
    G.add = (fz3)dlsym(h, "__gmpz_add");
    G.mul = (fz3)dlsym(h, "__gmpz_mul");
    G.tdiv_qr = (fqr)dlsym(h, "__gmpz_tdiv_qr");
    G.import = (fimp)dlsym(h, "__gmpz_import");
    G.export_ = (fexp)dlsym(h, "__gmpz_export");
    G.state = (G.init && G.clear && G.add && G.mul && G.tdiv_qr && G.import && G.export_) ? 1 : -1;
    return G.state > 0;
}

int og_available(void) { return og_load(); }

/* exports z into out[0..m), truncated mod B^m and zero-padded; buf holds at
 * least the value's digit count */
static void og_out(const og_mpz *z, uint32_t *buf, uint32_t *out, uint32_t m) {
    size_t cnt = 0;
    G.export_(buf, &cnt, -1, 4, 0, 0, z);
    const size_t k = cnt < m ? cnt : m;
    memcpy(out, buf, k * 4);
    if (k < m) memset(out + k, 0, (m - k) * 4);
}

/*
 * op 0: rs = (a + b) mod B^m_out        (a: ma digits, b: mb digits)
 * op 1: rs = (a * b) mod B^m_out
 * op 2: rs = floor(a / b) (m_out = ma digits), rs2 = a mod b (mb digits)
 * Returns 0, -1 if GMP is unavailable, -2 for a zero divisor (that
 * instance's outputs are zero, the rest complete).
 */
int og_batch(int op, uint32_t ma, uint32_t mb, uint32_t m_out, uint64_t n, const uint32_t *as, const uint32_t *bs,
             uint32_t *rs, uint32_t *rs2, int threads) {
    if (!og_load()) return -1;
    int bad = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(| : bad)
#endif
    {
        og_mpz a, b, r, s;
        G.init(&a);
        G.init(&b);
        G.init(&r);
        G.init(&s);
        uint32_t *buf = (uint32_t *)malloc(((size_t)ma + mb + 4) * 4);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t i = 0; i < (int64_t)n; i++) {
            G.import(&a, ma, -1, 4, 0, 0, as + (size_t)i * ma);
            G.import(&b, mb, -1, 4, 0, 0, bs + (size_t)i * mb);
            if (op == 0) {
                G.add(&r, &a, &b);
                og_out(&r, buf, rs + (size_t)i * m_out, m_out);
            } else if (op == 1) {
                G.mul(&r, &a, &b);
                og_out(&r, buf, rs + (size_t)i * m_out, m_out);
            } else {
                if (b.size == 0) {
                    memset(rs + (size_t)i * m_out, 0, (size_t)m_out * 4);
                    memset(rs2 + (size_t)i * mb, 0, (size_t)mb * 4);
                    bad |= 1;
                    continue;
                }
                G.tdiv_qr(&r, &s, &a, &b);
                og_out(&r, buf, rs + (size_t)i * m_out, m_out);
                og_out(&s, buf, rs2 + (size_t)i * mb, mb);
            }
        }
        free(buf);
        G.clear(&a);
        G.clear(&b);
        G.clear(&r);
        G.clear(&s);
    }
    return bad ? -2 : 0;
}

/*
 * Ragged batch, instance i in digits [off[i], off[i+1]) of as, bs, rs:
 * op 0: (a_i + b_i) mod B^{n_i}; op 1: (a_i * b_i) mod B^{n_i}.
 * Dynamic scheduling: the instances differ in cost by orders of magnitude.
 */
int og_batch_csr(int op, uint64_t n, const int64_t *off, const uint32_t *as, const uint32_t *bs, uint32_t *rs,
                 int threads) {
    if (!og_load()) return -1;
    uint32_t maxd = 0;
    for (uint64_t i = 0; i < n; i++)
        if ((uint32_t)(off[i + 1] - off[i]) > maxd) maxd = (uint32_t)(off[i + 1] - off[i]);
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        og_mpz a, b, r;
        G.init(&a);
        G.init(&b);
        G.init(&r);
        uint32_t *buf = (uint32_t *)malloc(((size_t)2 * maxd + 4) * 4);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t i = 0; i < (int64_t)n; i++) {
            const uint32_t m = (uint32_t)(off[i + 1] - off[i]);
            if (m == 0) continue;
            G.import(&a, m, -1, 4, 0, 0, as + off[i]);
            G.import(&b, m, -1, 4, 0, 0, bs + off[i]);
            if (op == 0) G.add(&r, &a, &b);
            else G.mul(&r, &a, &b);
            og_out(&r, buf, rs + off[i], m);
        }
        free(buf);
        G.clear(&a);
        G.clear(&b);
        G.clear(&r);
    }
    return 0;
}
