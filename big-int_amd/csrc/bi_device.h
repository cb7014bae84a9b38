// bi_device.h -- wave64 device primitives shared by every gfx950 kernel.
//
// Digit layout everywhere: little-endian uint32_t digits, an integer of m digits
// is m consecutive words, a batch is [n][m] instance-major (reference layout:
// cuda/main.cu:119, cuda/ker-helpers.cu.h:136,174-177).  A ragged batch keeps
// the same words back to back and an int64 offset table marks where each
// integer starts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bi {

constexpr int kWave = 64;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Wave-wide carry look-ahead (the associative 2-bit carry operator of
// cuda/ker-add.cu.h:10-23 evaluated on all 64 lanes at once).
//
// Lane i brings g_i = "its digits produce a carry when nothing comes in" and
// p_i = "an incoming carry runs through all its digits"; g_i and p_i are never
// both set.  With A = G|P and B = G (one bit per lane), A+B+cin is a 64-bit
// addition whose generate bits are A&B = G and propagate bits A^B = P, so the
// carry into lane i is bit i of (A+B+cin) ^ P and the carry out of lane 63 is
// the overflow of that addition.  G, P and the sum live in scalar registers:
// the whole scan costs two ballots and a handful of SALU ops.
__device__ __forceinline__ uint32_t wave_carry(bool g, bool p, uint32_t cin, uint32_t& cout) {
    const uint64_t G = ballot(g);
    const uint64_t P = ballot(p);
    const uint64_t A = G | P;
    const uint64_t S = A + G;
    const uint32_t c1 = S < A;
    const uint64_t S2 = S + (uint64_t)cin;
    const uint32_t c2 = S2 < S;
    cout = c1 | c2;
    const uint64_t C = S2 ^ P;
    return (uint32_t)(C >> lane_id()) & 1u;
}

// 64-lane inclusive max of a 32-bit value (used for most-significant-digit
// searches); returns the same value on every lane.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint32_t o = __shfl_xor(v, off, kWave);
        v = v > o ? v : o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v |= __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ uint32_t readfirst(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <bool NT>
__device__ __forceinline__ u32x4 ld4(const uint32_t* p) {
    if (NT) return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    return *reinterpret_cast<const u32x4*>(p);
}
template <bool NT>
__device__ __forceinline__ void st4(uint32_t* p, u32x4 v) {
    if (NT) __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(p));
    else *reinterpret_cast<u32x4*>(p) = v;
}

// 32x32->64 multiply-accumulate into a 64-bit column sum with an overflow
// counter: acc += a*b; cnt += carry-out.  v_mad_u64_u32 reports its own carry
// out, so one mad plus one addc covers the whole step.
__device__ __forceinline__ void mac(uint64_t& acc, uint32_t& cnt, uint32_t a, uint32_t b) {
    asm(
        "v_mad_u64_u32 %0, vcc, %2, %3, %0\n\t"
        "v_addc_co_u32 %1, vcc, 0, %1, vcc"
        : "+v"(acc), "+v"(cnt)
        : "v"(a), "v"(b)
        : "vcc");
}

// Four column sums at once: acc[c] += a * w_c, cnt[c] += carry-out.  The four
// mads put their carries in four SGPR pairs and the counts follow, so no
// count waits on the mad just before it.  (Not volatile: the outputs carry
// every effect, which lets the compiler wait only for the LDS reads each
// group needs and interleave the address updates.)
__device__ __forceinline__ void mac4(uint64_t* acc, uint32_t* cnt, uint32_t a, uint32_t w0, uint32_t w1, uint32_t w2,
                                     uint32_t w3) {
    uint64_t k0, k1, k2, k3;
    asm(
        "v_mad_u64_u32 %0, %8, %12, %13, %0\n\t"
        "v_mad_u64_u32 %1, %9, %12, %14, %1\n\t"
        "v_mad_u64_u32 %2, %10, %12, %15, %2\n\t"
        "v_mad_u64_u32 %3, %11, %12, %16, %3\n\t"
        "v_addc_co_u32_e64 %4, %8, 0, %4, %8\n\t"
        "v_addc_co_u32_e64 %5, %9, 0, %5, %9\n\t"
        "v_addc_co_u32_e64 %6, %10, 0, %6, %10\n\t"
        "v_addc_co_u32_e64 %7, %11, 0, %7, %11"
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(cnt[0]), "+v"(cnt[1]), "+v"(cnt[2]),
          "+v"(cnt[3]), "=&s"(k0), "=&s"(k1), "=&s"(k2), "=&s"(k3)
        : "v"(a), "v"(w0), "v"(w1), "v"(w2), "v"(w3));
}

// First index i in [lo, hi] with off[i] >= x, where off is non-decreasing;
// hi itself is treated as satisfying the bound (so hi is returned when no
// earlier index does).  One wave, 64-ary search; every lane returns the same
// value.  Only indices in [lo, hi] are read.
__device__ __forceinline__ uint64_t wave_lower_bound(const int64_t* __restrict__ off, uint64_t lo, uint64_t hi,
                                                     int64_t x) {
    const uint32_t lane = lane_id();
    while (hi - lo >= 64) {
        // 63*step >= hi-lo, so lane 63 always lands on (clamped) hi
        const uint64_t step = (hi - lo + 62) / 63;
        uint64_t idx = lo + (uint64_t)lane * step;
        if (idx > hi) idx = hi;
        const bool ge = (idx == hi) ? true : (off[idx] >= x);
        const uint32_t f = (uint32_t)__builtin_ctzll(ballot(ge));
        if (f == 0) return lo;
        const uint64_t nlo = lo + (uint64_t)(f - 1) * step + 1;
        const uint64_t cand = lo + (uint64_t)f * step;
        hi = cand < hi ? cand : hi;
        lo = nlo;
    }
    const uint64_t idx = lo + lane;
    const bool ge = (idx >= hi) ? true : (off[idx] >= x);
    return lo + (uint64_t)__builtin_ctzll(ballot(ge));
}

}  // namespace bi
