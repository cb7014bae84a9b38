// Host check of the prime-field arithmetic of big-int_amd/csrc/bi_ntt.h
// against 128-bit integer arithmetic: add, sub, mul on random and edge
// operands, the order of the roots of unity, inverses, and a cyclic
// convolution by transform against the direct sum.  Prints "bad 0" when all
// agree.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "bi_ntt.h"

using namespace bi;
typedef unsigned __int128 u128;

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main() {
    const uint64_t P = gl::P;
    long bad = 0;
    std::vector<uint64_t> edge = {0, 1, 2, gl::EPS - 1, gl::EPS, gl::EPS + 1, 1ull << 32, (1ull << 63),
                                  P - 2, P - 1, 0x7FFFFFFFFFFFFFFFull, 0xFFFFFFFF00000000ull};
    std::vector<uint64_t> vals = edge;
    for (int i = 0; i < 2000; i++) vals.push_back(rnd() % P);
    for (uint64_t a : vals)
        for (uint64_t b : vals) {
            const uint64_t s = (uint64_t)(((u128)a + b) % P);
            const uint64_t d = (uint64_t)(((u128)a + P - b) % P);
            const uint64_t m = (uint64_t)(((u128)a * b) % P);
            bad += gl::add(a, b) != s;
            bad += gl::sub(a, b) != d;
            bad += gl::mul(a, b) != m;
            uint64_t lo, hi;
            gl::mul128(a, b, lo, hi);
            bad += ((((u128)hi) << 64) | lo) != (u128)a * b;
        }
    // full 64-bit operands of mul128 (the reduction only sees canonical ones)
    for (int i = 0; i < 1000000; i++) {
        const uint64_t a = rnd(), b = rnd();
        uint64_t lo, hi;
        gl::mul128(a, b, lo, hi);
        bad += ((((u128)hi) << 64) | lo) != (u128)a * b;
        const u128 x = (((u128)rnd()) << 64) | rnd();
        bad += gl::reduce((uint64_t)x, (uint64_t)(x >> 64)) != (uint64_t)(x % P);
    }
    // roots of unity: w^(2^k) = 1 and w^(2^(k-1)) = p - 1
    for (int k = 1; k <= 32; k++) {
        const uint64_t w = gl::root_of_unity(k);
        bad += gl::pow(w, 1ull << k) != 1;
        bad += gl::pow(w, 1ull << (k - 1)) != P - 1;
    }
    for (uint64_t a : vals)
        if (a) bad += gl::mul(a, gl::inverse(a)) != 1;
    // cyclic convolution of length 64 by the transform
    {
        const int L = 64;
        const uint64_t w = gl::root_of_unity(6), wi = gl::inverse(w), li = gl::inverse(L);
        std::vector<uint64_t> x(L), y(L), X(L), Y(L), z(L), ref(L, 0);
        for (int i = 0; i < L; i++) {
            x[i] = rnd() & 0xFFFF;
            y[i] = rnd() & 0xFFFF;
        }
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) ref[(i + j) % L] += x[i] * y[j];
        for (int k = 0; k < L; k++) {
            X[k] = Y[k] = 0;
            for (int i = 0; i < L; i++) {
                const uint64_t t = gl::pow(w, (uint64_t)i * k % L);
                X[k] = gl::add(X[k], gl::mul(x[i], t));
                Y[k] = gl::add(Y[k], gl::mul(y[i], t));
            }
        }
        for (int i = 0; i < L; i++) {
            z[i] = 0;
            for (int k = 0; k < L; k++)
                z[i] = gl::add(z[i], gl::mul(gl::mul(X[k], Y[k]), gl::pow(wi, (uint64_t)i * k % L)));
            z[i] = gl::mul(z[i], li);
            bad += z[i] != ref[i];
        }
    }
    printf("bad %ld\n", bad);
    return 0;
}
