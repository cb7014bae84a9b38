// bi_ntt.h -- arithmetic in the prime field of p = 2^64 - 2^32 + 1, the field
// of the number-theoretic transform behind the long full products.  p - 1 =
// 2^32 * 3 * 5 * 17 * 257 * 65537, so the field holds 2^k-th roots of unity
// for every k <= 32; 7 generates the multiplicative group.  2^64 = 2^32 - 1
// and 2^96 = -1 (mod p), which makes the reduction of a 128-bit product a
// few 64-bit adds (no division, no Montgomery form).  Values are kept
// canonical (< p).  Host and device: the host build is what the unit check of
// these routines runs.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define BI_HD
#else
#define BI_HD __host__ __device__
#endif

namespace bi {
namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;   // 2^64 mod p
constexpr uint64_t GEN = 7;               // generator of the multiplicative group

// The corrections are selects, not branches: the transform runs them on
// every lane in lockstep.  The carries and borrows come from the additions
// themselves (the add-with-carry's carry out), not from separate compares.
BI_HD inline uint64_t add(uint64_t a, uint64_t b) {
    uint64_t s, t;
    // a + b >= 2^64: s + 2^64 = s + EPS (mod p); s >= p: s - p = s + EPS
    // (mod 2^64), and s + EPS then overflows.  The two exclude each other
    // (a + b < 2p), so either way the result is s + EPS mod 2^64.
    const bool c1 = __builtin_add_overflow(a, b, &s);
    const bool c2 = __builtin_add_overflow(s, EPS, &t);
    return (c1 | c2) ? t : s;
}

BI_HD inline uint64_t sub(uint64_t a, uint64_t b) {
    uint64_t d;
    // a < b: d = a - b + 2^64, and d - EPS = a - b + p
    const bool br = __builtin_sub_overflow(a, b, &d);
    return br ? d - EPS : d;
}

// 64 x 64 -> 128-bit product from four 32 x 32 -> 64-bit multiply-adds
BI_HD inline void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t m1 = (uint64_t)a0 * b1 + (p00 >> 32);          // < 2^64
    const uint64_t m2 = (uint64_t)a1 * b0 + (uint32_t)m1;         // < 2^64
    hi = (uint64_t)a1 * b1 + (m1 >> 32) + (m2 >> 32);
    lo = (m2 << 32) | (uint32_t)p00;
}

// (lo + 2^64 hi) mod p with hi = h1 2^32 + h0: lo - h1 + h0 (2^32 - 1)
BI_HD inline uint64_t reduce(uint64_t lo, uint64_t hi) {
    const uint64_t h1 = hi >> 32, h0 = hi & EPS;
    uint64_t d, r0, t;
    const bool br = __builtin_sub_overflow(lo, h1, &d);
    const uint64_t t0 = br ? d - EPS : d;          // borrow of 2^64 = EPS mod p
    const uint64_t t1 = (h0 << 32) - h0;           // h0 * EPS, no overflow
    const bool c = __builtin_add_overflow(t0, t1, &r0);
    const uint64_t r = c ? r0 + EPS : r0;          // carry of 2^64 = EPS mod p
    // r >= p: r - p = r + EPS (mod 2^64), which overflows exactly then
    const bool c2 = __builtin_add_overflow(r, EPS, &t);
    return c2 ? t : r;
}

BI_HD inline uint64_t mul(uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce(lo, hi);
}

BI_HD inline uint64_t pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mul(r, a);
        a = mul(a, a);
        e >>= 1;
    }
    return r;
}

// a primitive 2^k-th root of unity (k <= 32)
BI_HD inline uint64_t root_of_unity(int k) { return pow(GEN, (P - 1) >> k); }

BI_HD inline uint64_t inverse(uint64_t a) { return pow(a, P - 2); }

}  // namespace gl
}  // namespace bi
