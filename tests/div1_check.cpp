// Host check of the one-digit division helpers (bi_div1.h) against 64/128-bit
// integer division: quotient/remainder steps, modular products and sums over
// random and edge divisors (1, 2, powers of two, 2^32-1, ...), and the
// normalised-divisor variants (a whole long division against the plain one).
#include <cstdio>
#include <cstdlib>
#include <random>
#include "bi_div1.h"
using namespace bi;
int main() {
    std::mt19937_64 g(7);
    uint32_t ds[] = {1, 2, 3, 5, 7, 0x80000000u, 0x7FFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFEu, 65536, 65537, 1000000007u, 0x12345678u};
    long bad = 0, n = 0;
    for (int t = 0; t < 200000; t++) {
        uint32_t d = t < 13 * 1000 ? ds[t % 13] : (uint32_t)g();
        if (t % 7 == 0) d = (uint32_t)(g() >> (g() % 32));
        if (!d) d = 1;
        Div1 D = div1_make(d);
        for (int k = 0; k < 20; k++) {
            uint32_t x = (uint32_t)(g() % d), u = (uint32_t)g();
            if (k == 0) { x = d - 1; u = 0xFFFFFFFFu; }
            if (k == 1) { x = 0; u = 0; }
            uint32_t q, r;
            div_step(D, x, u, q, r);
            uint64_t num = ((uint64_t)x << 32) | u;
            if (q != num / d || r != num % d) bad++;
            uint32_t y = (uint32_t)(g() % d);
            if (k == 2) y = d - 1, x = d - 1;
            if (mulmod(D, x, y) != (uint32_t)(((unsigned __int128)x * y) % d)) bad++;
            if (addmod(D, x, y) != (uint32_t)(((uint64_t)x + y) % d)) bad++;
            // residues mod the normalised divisor
            const uint32_t xn = (uint32_t)(g() % D.dn), yn = k == 3 ? D.dn - 1 : (uint32_t)(g() % D.dn);
            if (mulmod_n(D, xn, yn) != (uint32_t)(((unsigned __int128)xn * yn) % D.dn)) bad++;
            if (addmod_n(D, xn, yn) != (uint32_t)(((uint64_t)xn + yn) % D.dn)) bad++;
            n += 5;
        }
        // a 9-digit dividend: schoolbook on the normalised digits (top digit
        // = the bits shifted out) against div_step on the plain digits
        uint32_t U[9];
        for (int j = 0; j < 9; j++) U[j] = (t % 5 == 0) ? 0xFFFFFFFFu : (uint32_t)g();
        uint32_t xa = 0, xb = D.s ? U[8] >> (32 - D.s) : 0u;
        for (int j = 8; j >= 0; j--) {
            uint32_t qa, qb;
            div_step(D, xa, U[j], qa, xa);
            div21(xb, norm_digit(U[j], j ? U[j - 1] : 0u, D.s), D.dn, D.v, qb, xb);
            if (qa != qb) bad++;
        }
        if ((xb >> D.s) != xa || (xb & ((1u << D.s) - 1u)) != 0) bad++;
        n += 10;
    }
    printf("checked %ld, bad %ld\n", n, bad);
    return bad != 0;
}
