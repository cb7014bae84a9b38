// bi_div1.h -- division of two-digit numbers by one invariant digit with a
// precomputed reciprocal (N. Moller and T. Granlund, "Improved division by
// invariant integers", IEEE Trans. Computers 60(2), 2011, Algorithm 4), and
// the modular products built on it.  Host and device: the host build is what
// the unit check of these routines runs.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define BI_HD
#else
#define BI_HD __host__ __device__
#endif

namespace bi {

struct Div1 {
    uint32_t d;    // divisor (nonzero)
    uint32_t dn;   // d << s, top bit set
    uint32_t v;    // floor((2^64 - 1) / dn) - 2^32
    uint32_t s;    // normalisation shift
};

BI_HD inline Div1 div1_make(uint32_t d) {
    Div1 r;
    r.d = d;
    const uint32_t s = (uint32_t)__builtin_clz(d);
    r.s = s;
    r.dn = d << s;
    r.v = (uint32_t)(0xFFFFFFFFFFFFFFFFull / r.dn - (1ull << 32));
    return r;
}

// (u1 * 2^32 + u0) / dn for u1 < dn: quotient and remainder (normalised).
BI_HD inline void div21(uint32_t u1, uint32_t u0, uint32_t dn, uint32_t v, uint32_t& q, uint32_t& r) {
    const uint64_t qq = (uint64_t)v * u1 + (((uint64_t)u1 << 32) | u0);
    uint32_t q1 = (uint32_t)(qq >> 32) + 1u;
    const uint32_t q0 = (uint32_t)qq;
    uint32_t rr = u0 - q1 * dn;
    if (rr > q0) {
        q1 -= 1u;
        rr += dn;
    }
    if (rr >= dn) {
        q1 += 1u;
        rr -= dn;
    }
    q = q1;
    r = rr;
}

// (x * 2^32 + u) / d for x < d: quotient digit and remainder
BI_HD inline void div_step(const Div1& D, uint32_t x, uint32_t u, uint32_t& q, uint32_t& r) {
    const uint32_t u1 = D.s ? (x << D.s) | (u >> (32 - D.s)) : x;
    const uint32_t u0 = u << D.s;
    uint32_t rn;
    div21(u1, u0, D.dn, D.v, q, rn);
    r = rn >> D.s;
}

// x * y mod d for x, y < d
BI_HD inline uint32_t mulmod(const Div1& D, uint32_t x, uint32_t y) {
    const uint64_t p = ((uint64_t)x * y) << D.s;   // < d * dn < 2^64
    uint32_t q, rn;
    div21((uint32_t)(p >> 32), (uint32_t)p, D.dn, D.v, q, rn);
    return rn >> D.s;
}

BI_HD inline uint32_t addmod(const Div1& D, uint32_t x, uint32_t y) {
    const uint64_t t = (uint64_t)x + y;
    return (uint32_t)(t >= D.d ? t - D.d : t);
}

// The same routines on the normalised divisor dn = d * 2^s, for dividing the
// normalised dividend N = U * 2^s (digit j of N = bits [32j - s, 32j + 32 - s)
// of U, see norm_digit): N / dn = U / d, N mod dn = (U mod d) * 2^s, and every
// step is a div21 with no shifts.  Residues here are mod dn.
BI_HD inline uint32_t norm_digit(uint32_t u, uint32_t below, uint32_t s) {
    return s ? (uint32_t)((((uint64_t)u << 32) | below) >> (32 - s)) : u;
}

// x * y mod dn for x, y < dn
BI_HD inline uint32_t mulmod_n(const Div1& D, uint32_t x, uint32_t y) {
    const uint64_t p = (uint64_t)x * y;   // < dn^2: high word < dn
    uint32_t q, r;
    div21((uint32_t)(p >> 32), (uint32_t)p, D.dn, D.v, q, r);
    return r;
}

BI_HD inline uint32_t addmod_n(const Div1& D, uint32_t x, uint32_t y) {
    const uint64_t t = (uint64_t)x + y;
    return (uint32_t)(t >= D.dn ? t - D.dn : t);
}

}  // namespace bi
