// bi_addcore.h -- the segmented carry-scan span shared by the add/sub
// kernels and the persistent ragged kernel.
//
// Reference semantics: rs = (as + bs) mod B^m per instance with the carry out
// dropped (cuda/ker-add.cu.h:57-66, baddKer1/2/3 at :50-266); subtraction
// follows the borrow formulation of futhark/sub.fut:14-26.
//
// A batch, uniform or ragged, is one flat array of digits.  A wave owns a span
// of whole instances (the standalone ragged add: a fixed chunk of the flat
// array, add_flat_chunk) and walks it in 256-digit tiles (4 digits per lane, one
// 16-byte load per operand per lane, fully coalesced).  Inside a tile every
// lane adds its 4 digits with a local carry chain, reports generate/propagate
// bits, and the 64-lane carry look-ahead (bi::wave_carry) resolves all carries
// in scalar registers.  The carry out of lane 63 feeds the next tile; instance
// starts inside a tile cut the chain.  Every digit is read once and written
// once: 3 words per digit through HBM.
#pragma once
#include "bi_device.h"

namespace bi {


// Per-lane work on 4 digits.  fl: bit j set when digit j starts an instance.
// SUB computes a - b with borrows in place of carries.
template <bool SUB>
__device__ __forceinline__ void lane_pre(const u32x4 a, const u32x4 b, uint32_t fl, u32x4& s, bool& g, bool& p) {
    uint32_t c = 0;
    uint32_t all = SUB ? 0u : 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c &= ~(fl >> j) & 1u;
        uint32_t r;
        if (!SUB) {
            const uint64_t t = (uint64_t)a[j] + b[j] + c;
            r = (uint32_t)t;
            c = (uint32_t)(t >> 32);
            all &= r;
        } else {
            const uint64_t t = (uint64_t)a[j] - b[j] - c;
            r = (uint32_t)t;
            c = (uint32_t)(t >> 63);
            all |= r;
        }
        s[j] = r;
    }
    g = c != 0;
    p = (fl == 0) && (SUB ? (all == 0u) : (all == 0xFFFFFFFFu));
}

template <bool SUB>
__device__ __forceinline__ u32x4 lane_post(u32x4 s, uint32_t fl, uint32_t e) {
    u32x4 r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        e &= ~(fl >> j) & 1u;
        const uint32_t x = s[j];
        r[j] = SUB ? x - e : x + e;
        e &= SUB ? (uint32_t)(x == 0u) : (uint32_t)(x == 0xFFFFFFFFu);
    }
    return r;
}

// Loads digits [f, f+4) (zero beyond L).  VEC: 16-byte aligned vector load.
template <bool VEC, bool NT>
__device__ __forceinline__ u32x4 load_grp(const uint32_t* __restrict__ x, uint64_t f, uint64_t L) {
    if (VEC && f + 4 <= L) return ld4<NT>(x + f);
    u32x4 r = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (f + j < L) r[j] = x[f + j];
    return r;
}

template <bool VEC, bool NT>
__device__ __forceinline__ void store_grp(uint32_t* __restrict__ x, uint64_t f, u32x4 v, uint64_t s0, uint64_t s1) {
    if (VEC && f >= s0 && f + 4 <= s1) {
        st4<NT>(x + f, v);
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (f + j >= s0 && f + j < s1) x[f + j] = v[j];
}

// Segment-start flags of the 4 digits at flat index f for a uniform size m:
// digit f+j starts an instance when (f+j-s0) % m == 0 (s0 is an instance start).
__device__ __forceinline__ uint32_t uniform_flags(uint64_t f, uint64_t s0, uint32_t m) {
    if (f + 4 <= s0) return 0u;
    uint32_t fl = 0;
    if (f >= s0) {
        uint32_t r = (uint32_t)(f - s0) % m;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            fl |= (r == 0u ? 1u : 0u) << j;
            r = (r + 1 == m) ? 0u : r + 1;
        }
    } else {
        // the tile straddles s0: only digits at/after s0 can start instances
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (f + j >= s0) fl |= (((uint32_t)(f + j - s0) % m) == 0u ? 1u : 0u) << j;
        }
    }
    return fl;
}

// CSR segment flags for the tile starting at flat index tb.  The span's
// instance starts are held one per lane (`has`: this lane holds one).  The
// lanes whose start falls in [tb, tb+256) are balloted, and for each of them
// (lowest first) the lane owning that digit sets its bit: every start lies in
// one tile, so the loop runs at most once per instance of the span.
__device__ __forceinline__ uint32_t csr_flags(uint64_t start, bool has, uint64_t tb) {
    const uint32_t lane = lane_id();
    const uint64_t st = start - tb;
    uint64_t hit = __ballot(has && st < 256);
    uint32_t fl = 0;
    while (hit) {
        const uint32_t j = (uint32_t)__builtin_ctzll(hit);
        hit &= hit - 1;
        const uint32_t d = __builtin_amdgcn_readlane((uint32_t)st, j);
        if ((d >> 2) == lane) fl |= 1u << (d & 3);
    }
    return fl;
}

// a < b for one instance of m digits, decided by the whole wave (uniform
// result): the top 64 digits first, then 256-digit tiles downward until one
// differs (random operands: the first load decides).
__device__ __forceinline__ bool wave_less(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                          uint32_t m) {
    const uint32_t lane = lane_id();
    const uint32_t top = m > kWave ? m - kWave : 0;
    {
        const uint32_t k = top + lane;
        bool ne = false, lt = false;
        if (k < m) {
            const uint32_t x = a[k], y = b[k];
            ne = x != y;
            lt = x < y;
        }
        const uint64_t d = __ballot(ne);
        if (d) return (__ballot(lt) >> (63 - __builtin_clzll(d))) & 1u;
    }
    for (int64_t tb = top ? (int64_t)((top - 1) & ~255u) : -1; tb >= 0; tb -= 256) {
        bool ne = false, lt = false;
#pragma unroll
        for (int j = 0; j < 4; j++) {   // ascending: the higher digit decides
            const uint32_t k = (uint32_t)tb + 4 * lane + j;
            if (k < top) {
                const uint32_t x = a[k], y = b[k];
                if (x != y) {
                    ne = true;
                    lt = x < y;
                }
            }
        }
        const uint64_t d = __ballot(ne);
        if (d) return (__ballot(lt) >> (63 - __builtin_clzll(d))) & 1u;
    }
    return false;
}

// One wave per span of whole instances, instances [w*ipw, min((w+1)*ipw, n));
// WPB waves per workgroup; U tiles of loads in flight per wave.
//   uniform: instances of size m.
//   CSR:     ipw <= 63; lane j holds off[i0+j] (the span's starts and its end),
//            read once, so the span needs no search and its tiles no further
//            offset loads.
//   |a - b| (SUB, uniform, `order` given): instances flagged in `order`
//            subtract a from b.  CMP (ipw <= 63): the wave decides a < b for
//            its own instances first (wave_less), keeps the flags as a mask
//            and writes them to `order`; otherwise `order` is read.
template <bool SUB, bool CSR, bool VEC, bool NT, int U, bool CMP = false>
__device__ __forceinline__ void addsub_span(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                            uint32_t* __restrict__ rs, uint64_t n, uint32_t m, uint64_t ipw,
                                            const int64_t* __restrict__ off, uint64_t w,
                                            uint8_t* __restrict__ order) {
    const uint32_t lane = lane_id();

    const uint64_t i0 = w * ipw;
    if (i0 >= n) return;
    const uint64_t i1 = i0 + ipw < n ? i0 + ipw : n;
    uint64_t s0, s1, start = 0;
    bool has_start = false;
    if (!CSR) {
        s0 = i0 * m;
        s1 = i1 * m;
    } else {
        const uint32_t k = (uint32_t)(i1 - i0);
        start = lane <= k ? (uint64_t)off[i0 + lane] : 0ull;
        has_start = lane < k;
        const uint32_t lo = (uint32_t)start, hi = (uint32_t)(start >> 32);
        s0 = ((uint64_t)__builtin_amdgcn_readfirstlane(hi) << 32) | __builtin_amdgcn_readfirstlane(lo);
        s1 = ((uint64_t)__builtin_amdgcn_readlane(hi, k) << 32) | __builtin_amdgcn_readlane(lo, k);
    }
    if (s1 <= s0) return;
    uint64_t omask = 0;   // CMP: bit q = (a < b) for instance i0 + q
    if (CMP) {
        for (uint32_t q = 0; q < (uint32_t)(i1 - i0); q++)
            omask |= (uint64_t)wave_less(as + (i0 + q) * m, bs + (i0 + q) * m, m) << q;
        if (lane < i1 - i0) order[i0 + lane] = (uint8_t)((omask >> lane) & 1u);
    }

    uint32_t carry = 0;
    const uint64_t tfirst = s0 & ~(uint64_t)255;
    for (uint64_t tb = tfirst; tb < s1; tb += 256 * U) {
        u32x4 av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t f = tb + 256 * u + 4 * lane;
            if (tb + 256 * u < s1) {
                // digits outside [s0, s1) belong to the spans around (read once)
                const bool mine = f + 4 > s0;
                av[u] = mine ? load_grp<VEC, NT>(as, f, s1) : u32x4{0u, 0u, 0u, 0u};
                bv[u] = mine ? load_grp<VEC, NT>(bs, f, s1) : u32x4{0u, 0u, 0u, 0u};
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t t = tb + 256 * u;
            if (t >= s1) break;
            const uint64_t f = t + 4 * lane;
            const uint32_t fl = CSR ? csr_flags(start, has_start, t) : uniform_flags(f, s0, m);
            if (SUB && !CSR && order && f + 4 > s0 && f < s1) {
                // |a - b|: instances flagged in `order` subtract a from b.  The
                // lane's first digit of this span, d0, lies in instance i0+q at
                // offset r; a flag is loaded once per instance the lane touches
                // (at most two for m >= 4).
                const uint64_t d0 = f > s0 ? f : s0;
                const uint32_t rel = (uint32_t)(d0 - s0);   // < ipw*m digits of one span
                const uint32_t q = rel / m;
                uint32_t r = rel - q * m;
                uint64_t inst = i0 + q;
                bool o = CMP ? ((omask >> q) & 1u) != 0 : order[inst] != 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (f + j >= d0 && f + j < s1) {
                        if (o) {
                            const uint32_t x = av[u][j];
                            av[u][j] = bv[u][j];
                            bv[u][j] = x;
                        }
                        if (++r == m) {
                            r = 0;
                            inst++;
                            o = CMP ? ((omask >> (inst - i0)) & 1u) != 0 : (inst < n && order[inst] != 0);
                        }
                    }
                }
            }
            u32x4 s;
            bool g, p;
            lane_pre<SUB>(av[u], bv[u], fl, s, g, p);
            uint32_t cout;
            const uint32_t e = wave_carry(g, p, carry, cout);
            carry = cout;
            store_grp<VEC, NT>(rs, f, lane_post<SUB>(s, fl, e), s0, s1);
        }
    }
}

// Chunk w of a uniform batch whose instances of m digits are cut into cpi =
// ceil(m / D) chunks of D digits (D a multiple of 256): digits [s0, s1) of
// one instance added with carry-in 0 (a segment flag at s0 cuts whatever
// comes from below).  Returns the chunk's state for the carry pass between
// chunks: 1 if a carry leaves the chunk, 2 if every digit of the chunk is
// all-ones (an incoming carry runs through it), 0 if neither.  The two
// exclude each other: with carry-in 0 an all-ones chunk generates nothing.
template <bool VEC, bool NT, int U>
__device__ __forceinline__ uint32_t add_chunk(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                              uint32_t* __restrict__ rs, uint32_t m, uint32_t D, uint32_t cpi,
                                              uint64_t w) {
    const uint32_t lane = lane_id();
    const uint64_t i = w / cpi;
    const uint32_t c = (uint32_t)(w - i * cpi);
    const uint64_t s0 = i * m + (uint64_t)c * D;
    const uint64_t s1 = i * m + ((uint64_t)c * D + D < m ? (uint64_t)c * D + D : (uint64_t)m);
    uint32_t carry = 0, out = 0;
    bool allones = true;
    for (uint64_t tb = s0 & ~(uint64_t)255; tb < s1; tb += 256 * U) {
        u32x4 av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t f = tb + 256 * u + 4 * lane;
            if (tb + 256 * u < s1) {
                const bool mine = f + 4 > s0;
                av[u] = mine ? load_grp<VEC, NT>(as, f, s1) : u32x4{0u, 0u, 0u, 0u};
                bv[u] = mine ? load_grp<VEC, NT>(bs, f, s1) : u32x4{0u, 0u, 0u, 0u};
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t t = tb + 256 * u;
            if (t >= s1) break;
            const uint64_t f = t + 4 * lane;
            const uint32_t fl = (f <= s0 && s0 < f + 4) ? 1u << (uint32_t)(s0 - f) : 0u;
            u32x4 sm;
            bool g, p;
            lane_pre<false>(av[u], bv[u], fl, sm, g, p);
            uint32_t cout;
            const uint32_t e = wave_carry(g, p, carry, cout);
            carry = cout;
            const u32x4 r = lane_post<false>(sm, fl, e);
            store_grp<VEC, NT>(rs, f, r, s0, s1);
            bool mine1 = true;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (f + j >= s0 && f + j < s1) mine1 &= r[j] == 0xFFFFFFFFu;
            allones &= ballot(!mine1) == 0;
            // the carry out of digit s1-1: the sum at s1 (whose operands read as zero), or the tile's carry out
            if (s1 < t + 256) {
                const uint32_t k = (uint32_t)(s1 - t);
                const uint32_t x = __builtin_amdgcn_readlane((k & 3) == 0 ? r[0] : (k & 3) == 1 ? r[1]
                                                             : (k & 3) == 2 ? r[2] : r[3], k >> 2);
                out = x;
            } else if (s1 == t + 256) {
                out = cout;
            }
        }
    }
    return allones ? 2u : (out & 1u);
}

// 64-bit readlane (a wave-uniform result).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l) {
    return ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), l) << 32) |
           __builtin_amdgcn_readlane((uint32_t)v, l);
}

// Ragged add over fixed chunks of the flat digit array: wave w adds digits
// [w D, (w + 1) D) (D a multiple of 256) whichever instances own them, so
// every wave has the same work and every tile is one whole 256-digit group
// of 16-byte loads however the instance sizes vary (a span of whole
// instances reads the tiles around its ends twice and varies in length).
//  * The chunk's first tiles are loaded before anything else.
//  * The instance that holds the chunk's first digit s0: a 64-ary search of
//    the offsets (two levels and a last one for 2^17 instances; the last
//    load also brings the next 64 instance starts into the lanes, later
//    starts 64 at a time as the tiles reach them).
//  * The carry into s0, when s0 continues an instance, is recomputed from
//    the digits below it (the chunk before owns them): the 16 digits below
//    (one 64-byte segment an operand, one a lane in lanes 48..63) decide it
//    unless all 16 sums are all-ones, then 256-digit windows further down,
//    to the instance's start at the latest.  No wave waits for another.
// off: n + 1 non-decreasing offsets, total = off[n]; digits before off[0]
// belong to no instance and are neither read for a result nor written.
template <bool NT, int U>
__device__ __forceinline__ void add_flat_chunk(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                               uint32_t* __restrict__ rs, const int64_t* __restrict__ off, uint64_t n,
                                               uint64_t total, uint32_t D, uint64_t w) {
    const uint32_t lane = lane_id();
    const uint64_t c0 = w * D;
    const uint64_t c1 = c0 + D < total ? c0 + D : total;
    // the offset search's first probes go out before the data: vector loads
    // complete in issue order, so a probe issued after the tiles would wait
    // for all of them
    const uint64_t step1 = (n + kWave - 1) / kWave;
    uint64_t v1 = ~0ull;
    if (n > (uint64_t)kWave && (uint64_t)lane * step1 < n) v1 = (uint64_t)off[(uint64_t)lane * step1];
    // the first tiles and the 16 digits below the chunk go out before off[0]
    // is known: they lie in [0, total) whatever it is (digits below off[0]
    // are read but never used)
    u32x4 av[U], bv[U];
    auto load = [&](uint64_t tb) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t f = tb + 256 * u + 4 * lane;
            if (tb + 256 * u < c1) {
                av[u] = load_grp<true, NT>(as, f, c1);
                bv[u] = load_grp<true, NT>(bs, f, c1);
            }
        }
    };
    load(c0);
    // the 16 digits below the chunk in lanes 48..63 (one 64-byte segment an
    // operand; c0 >= D >= 256)
    uint32_t xa = 0, xb = 0;
    if (c0 && lane >= kWave - 16) {
        xa = as[c0 - kWave + lane];
        xb = bs[c0 - kWave + lane];
    }
    const uint64_t off0 = (uint64_t)off[0];
    if (c1 <= off0) return;
    const uint64_t s0 = c0 > off0 ? c0 : off0;
    const uint64_t tfirst = c0;
    // j = the last index with off[j] <= s0: off[lo] <= s0 < off[hi] throughout
    // (off[0] = off0 <= s0 < c1 <= total = off[n]); sl = off[lo]
    uint64_t lo = 0, hi = n, sl = off0;
    bool first = true;   // its probes (lo = 0, hi = n, step1) are v1
    while (hi - lo > (uint64_t)kWave) {
        const uint64_t step = (hi - lo + kWave - 1) / kWave;
        const uint64_t p = lo + (uint64_t)lane * step;
        const uint64_t v = first ? v1 : (p < hi ? (uint64_t)off[p] : ~0ull);
        first = false;
        const uint64_t le = ballot(v <= s0);   // lane 0 (p = lo) always
        const uint32_t L = 63u - (uint32_t)__builtin_clzll(le);
        sl = readlane64(v, L);
        lo += (uint64_t)L * step;
        hi = lo + step < hi ? lo + step : hi;
    }
    // the last level: lanes read off[lo + 1 + lane] (up to off[n]); those <=
    // s0 are a prefix of the lanes.  The same values are the first window of
    // instance starts (wb = lo + 1); starts below the tile are ignored.
    uint64_t wb = lo + 1;
    uint64_t start = wb + lane <= n ? (uint64_t)off[wb + lane] : ~0ull;
    const uint32_t cnt = (uint32_t)__builtin_popcountll(ballot(start <= s0));
    const uint64_t sj = cnt ? readlane64(start, cnt - 1) : sl;   // the start of the instance holding s0
    // s0 starts an instance: its flag is set here (off[j] may lie outside the
    // window when the window begins past j)
    const bool head = sj == s0;
    uint32_t carry = 0;
    if (sj < s0) {   // s0 = c0 continues instance j: the carry from its digits below
        const uint64_t d = c0 - kWave + lane;
        const uint64_t t = (uint64_t)xa + xb;
        const bool on = lane >= kWave - 16;
        const bool g = on && d >= sj && (t >> 32) != 0;
        const bool p = on && d > sj && (uint32_t)t == 0xFFFFFFFFu;
        (void)wave_carry(g, p, 0u, carry);
        if ((ballot(p) >> (kWave - 16)) == 0xFFFFull) {
            // all 16 sums all-ones: 256-digit windows further down; the one
            // holding sj has a lane that propagates nothing, so the loop ends there
            for (uint64_t tb = c0 - 16 - 256;; tb -= 256) {
                const uint64_t f = tb + 4 * lane;   // tb may wrap below 0: every read lies at or above sj
                u32x4 a4 = {0u, 0u, 0u, 0u}, b4 = {0u, 0u, 0u, 0u};
                if ((int64_t)f + 4 > (int64_t)sj) {
                    a4 = ld4<false>(as + f);
                    b4 = ld4<false>(bs + f);
                }
                const uint32_t fl = ((int64_t)f <= (int64_t)sj && (int64_t)sj < (int64_t)f + 4)
                                        ? 1u << (uint32_t)(sj - f) : 0u;
                u32x4 sm;
                bool g4, p4;
                lane_pre<false>(a4, b4, fl, sm, g4, p4);
                uint32_t cw;
                (void)wave_carry(g4, p4, 0u, cw);
                if (ballot(p4) != ~0ull) {
                    carry = cw;
                    break;
                }
            }
        }
    }
    for (uint64_t tb = tfirst; tb < c1; tb += 256 * U) {
        if (tb != tfirst) load(tb);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t t = tb + 256 * u;
            if (t >= c1) break;
            const uint64_t f = t + 4 * lane;
            // the tile's instance starts, from as many windows as reach it
            uint32_t fl = (head && f <= s0 && s0 < f + 4) ? 1u << (uint32_t)(s0 - f) : 0u;
            for (;;) {
                fl |= csr_flags(start, start != ~0ull, t);
                if (readlane64(start, kWave - 1) >= t + 256 || wb + kWave > n) break;
                wb += kWave;
                start = wb + lane <= n ? (uint64_t)off[wb + lane] : ~0ull;
            }
            u32x4 sm;
            bool g, p;
            lane_pre<false>(av[u], bv[u], fl, sm, g, p);
            uint32_t cout;
            const uint32_t e = wave_carry(g, p, carry, cout);
            carry = cout;
            store_grp<true, NT>(rs, f, lane_post<false>(sm, fl, e), s0, c1);
        }
    }
}

}  // namespace bi
