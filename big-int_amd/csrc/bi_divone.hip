// bi_divone.hip -- division of a small batch in ONE launch: one workgroup (256
// or 1024 lanes) per instance runs the whole Newton pipeline of bi_div.hip with
// every working buffer in LDS, its own precision schedule and no host round
// trip.
//
// Reference: prototypes/div.c:352-467 (shinv, refine, the delta fix),
// thesis/sections/division.tex:140-405.  The arithmetic is the batched
// pipeline's (bi_div.hip), per instance instead of per batch:
//
//   NV = max(prec(v), 3), s = NV - prec(v): V = v B^s has its top digit at NV-1,
//   U = u B^s (NU = prec(u) + s digits);  P = max(prec(u) - prec(v) + 1, 1) + 1;
//   the inverse's precision T = P + 2, or, for long quotients (in two
//   halves, below), T = max(P - k, k + 2) + 2 with k = (P - 1) / 2;
//   p0 = the schedule's largest precision of at most 31 digits (T itself
//   when T <= 31), X = floor(B^(2 p0) / (Y + 1)), Y = V's top p0 digits;
//   Newton steps p -> p2 from p0 (the schedule of bi_div.hip, from T down):
//     E  = B^(NV+p) - V X  = (-V X) mod B^(NV+2)          (low product, negated;
//          V's digits from cw = c - p - 1 up only, E / B^c then <= 1 unit high)
//     X' = X B^(p2-p) + digits [d, d+p2+2) of X (E / B^c)  (top columns only),
//          minus 1 when cw > 0 (cw = 0 where the window would save little)
//   Q  = digits [NV+T, NV+T+P) of U W, top columns only;
//   R  = (U - Q V) mod B^(NV+1); twice: if R >= V then Q += 1, R -= V;
//   q = Q, r = R / B^s.
// Two halves (P >= kOneSplitMinP256 / kOneSplitMinP1024): q1 and r1 as above from U2 = floor(U / B^k), then
// q0 and R from U3 = r1 B^k + (U mod B^k) < V B^k (q0 < B^k), q = q1 B^k +
// q0: the inverse needs half the precision (no last Newton step) and each
// quotient product is half as long.
// The start X comes from a long division on one wave (wave_recip_top): the
// batched pipeline starts at 3 digits, and its first steps (4, 5, 7, 11, 19
// digits) cost a workgroup about 5 us each here, for almost no arithmetic.
//
// Why one launch: the batched pipeline is about 60 dependent kernels and two
// host reads per call.  At small counts (a serving round of 1-256 divisions)
// every one of those kernels is short and the call's latency is their chain;
// here the chain is a sequence of workgroup barriers, and each instance runs
// only as many Newton steps as its own precision needs.
//
// Products (blk_prod): the column sums of columns [klo, klo + 4G) of A B in
// 4-column groups, in tiles of 64 groups by 8-128 rows of the shorter
// operand that whole waves draw from an LDS counter.  A lane walks its group's row blocks of the tile: 8
// digits of A (one broadcast 16-byte pair, the same rows for every lane)
// against a 12-digit window of B (three 16-byte reads, consecutive across the
// lanes), 32 multiply-adds into four 64-bit sums with overflow counts
// (bi::mac4), then adds its sums to the group's columns with LDS atomics
// (64-bit add with return, overflow into the 32-bit count).  Operands are
// read in place: every buffer has 16 zero words before and after it and its
// digits above the live length are zero.
//
// Carry pass (blk_digits): digit k of sum_k S_k B^k (+ an added operand's
// digit x_k, + B - 1 from the output window's first digit up when one is
// subtracted there), S_k = (lo64, cnt), is t_k = lo32(S_k) + hi32(S_(k-1)) +
// cnt(S_(k-2)) + x_k (+ B - 1) < 5 B, resolved in three steps: t = a + e B
// (e <= 4), y_k = a_k + e_(k-1) = b + f B (f <= 1, and b <= 3 when f = 1),
// z_k = b_k + f_(k-1) <= B; z_k = B needs b_k = B - 1, which rules out a carry into k, so
// one binary carry look-ahead over z (a wave's ballots, then the 16 waves'
// aggregates through LDS) gives the digits.  The Newton update X' = X B^(p2-p)
// + Y enters as the added operand, and the negated products (E, and R = U -
// Q V) run a second look-ahead in the same LDS round.
#include <algorithm>

#include "../../include/bigint.h"
#include "bi_addcore.h"
#include "bi_ctx.h"
#include "bi_div1.h"

namespace bi {

constexpr int kOneWaves = 16;   // most waves a workgroup (1024 lanes); LDS slot arrays are sized for them
constexpr uint32_t kOneGuard = 16;   // zero words before and after every buffer
constexpr uint32_t kOneRows = 8;     // digits of A per block step
constexpr uint32_t kOneStripe = 128; // rows of a product tile
constexpr uint32_t kOneLdsMax = 160 * 1024;
constexpr uint32_t kOneSmallDigits = 128;   // operands up to this many digits: 256-lane workgroups
constexpr uint32_t kOneDirect = 31;         // precision of the directly computed start (lanes: 2 p0 + 1 <= 64)
constexpr uint32_t kOneWindowMin = 8192;    // digit products a step's divisor window must save (its staging costs a pass)
// Quotients of at least this many digits + 1 come in two halves: with 256
// lanes (four workgroups a CU, product-bound) from 64 on; with 1024 lanes
// (one a CU, barrier-bound: the second quotient pass costs about as many
// barriers as the Newton step it saves) from 768 on.  One-launch time, n =
// 16 on 1024 lanes, without / with the halves (us): 256/128 digits 60.7 /
// 64.0, 768/384 82.1 / 84.9, 1024/512 95.0 / 95.7, 1536/768 111.1 / 108.6,
// 4096/2048 262 / 225; n = 1024 on 256 lanes: 128/64 61.4 / 58.7, 256/128
// 79.9 / 75.0, 1024/512 172.5 / 150.4.
constexpr uint32_t kOneSplitMinP256 = 64;
constexpr uint32_t kOneSplitMinP1024 = 768;

__host__ __device__ inline uint32_t one_r4(uint32_t x) { return (x + 3) & ~3u; }

// LDS layout (word offsets) of one instance with divisor length NV, precision
// P and numerator length NU.  E holds R after the Newton steps; Q goes into
// the X buffer W does not use.  Every buffer has kOneGuard zero words on
// either side.
struct OneLayout {
    uint32_t V, E, X0, X1, U, lo, cnt, ctr, red, scan, K, total;
};

__host__ __device__ inline OneLayout one_layout(uint32_t NV, uint32_t P, uint32_t NU) {
    OneLayout L;
    uint32_t o = kOneGuard;
    const uint32_t XC = P + 4;   // PP + 2 digits
    L.V = o;
    o += one_r4(NV + 4) + kOneGuard;
    L.E = o;
    o += one_r4(NV + 4) + kOneGuard;
    L.X0 = o;
    o += one_r4(XC + 4) + kOneGuard;
    L.X1 = o;
    o += one_r4(XC + 4) + kOneGuard;
    L.U = o;
    o += one_r4(NU + 4) + kOneGuard;
    // every product has at most max(NV, P) + 12 columns (rounded to groups)
    L.K = one_r4((NV > P ? NV : P) + 24);
    L.lo = (o + 1) & ~1u;   // 8-byte words
    o = L.lo + 2 * L.K;
    L.cnt = o;
    o += L.K;
    L.ctr = o;
    o += 4;
    L.red = o;
    o += 2 * kOneWaves;
    L.scan = o;
    o += 2 * kOneWaves;
    L.total = one_r4(o);
    return L;
}

// The largest layout any instance of a (mu, mv) batch can take: every divisor
// length NV (3 covers prec(v) <= 3) with the longest precision it allows.
static uint32_t one_lds_words(uint32_t mu, uint32_t mv) {
    uint32_t best = 0;
    const uint32_t nvmax = std::max<uint32_t>(mv, 3);
    for (uint32_t NV = 3; NV <= nvmax; NV++) {
        // NV = 3: prec(v) may be 1 (P up to mu + 1, s = 2); otherwise prec(v) = NV
        const uint32_t sv = NV == 3 ? 1 : NV;
        const uint32_t P = (mu + 1 > sv ? mu + 1 - sv : 1) + 1;
        const uint32_t NU = mu + (NV - sv);
        best = std::max(best, one_layout(NV, P, NU).total);
    }
    return best;
}

bool divone_fits(uint32_t mu, uint32_t mv) {
    return mu >= 1 && mv >= 1 && mu <= 8192 && mv <= 8192 && (size_t)one_lds_words(mu, mv) * 4 <= kOneLdsMax;
}

__device__ __forceinline__ uint32_t blk_wave() { return threadIdx.x / kWave; }

// Block-wide maximum; red: 2 * kOneWaves words, ph alternates the halves.
template <int NT>
__device__ __forceinline__ uint32_t blk_max(uint32_t x, uint32_t* red, uint32_t& ph) {
    x = wave_max_u32(x);
    uint32_t* sl = red + (ph & 1u) * kOneWaves;
    ph++;
    if (lane_id() == 0) sl[blk_wave()] = x;
    __syncthreads();
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < NT / kWave; j++) m = sl[j] > m ? sl[j] : m;
    return m;
}

// Block-wide carry look-ahead: the carry into this lane given cin into lane 0
// of wave 0; tout = the carry out of the last lane.  Each wave's aggregate
// (carry out for carry-in 0, "every lane propagates") meets in LDS; scan:
// 2 * kOneWaves words, halves alternating (a half is rewritten only after
// the next call's barrier, by which time every lane has read it).
template <int NT>
__device__ __forceinline__ uint32_t blk_carry(bool g, bool p, uint32_t cin, uint32_t* scan, uint32_t& ph,
                                              uint32_t& tout) {
    uint32_t wout;
    (void)wave_carry(g, p, 0u, wout);
    const bool allp = ballot(p) == ~0ull;
    uint32_t* sl = scan + (ph & 1u) * kOneWaves;
    ph++;
    const uint32_t w = blk_wave();
    if (lane_id() == 0) sl[w] = wout | (allp ? 2u : 0u);
    __syncthreads();
    uint32_t c = cin, cw = 0;
#pragma unroll
    for (int j = 0; j < NT / kWave; j++) {
        if (j == (int)w) cw = c;
        const uint32_t s = sl[j];
        c = (s & 1u) | ((s >> 1) & c);
    }
    tout = c;
    uint32_t dummy;
    return wave_carry(g, p, cw, dummy);
}

// r = X B^dx +- Y + cin over len digits (mod B^len).  r may be X (dx = 0) or
// Y: every lane reads its own digits before the scan's barrier and writes
// them after it.
template <bool SUB, int NT>
__device__ void blk_addsub(const uint32_t* X, uint32_t xlen, uint32_t dx, const uint32_t* Y, uint32_t ylen,
                           uint32_t cin, uint32_t* r, uint32_t len, uint32_t* scan, uint32_t& ph) {
    uint32_t c = cin;
    for (uint32_t base = 0; base < len; base += 4 * NT) {
        const uint32_t k = base + 4 * threadIdx.x;
        u32x4 x, y;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t kk = k + j;
            x[j] = (kk >= dx && kk - dx < xlen) ? X[kk - dx] : 0u;
            y[j] = kk < ylen ? Y[kk] : 0u;
        }
        u32x4 s;
        bool g, p;
        lane_pre<SUB>(x, y, 0u, s, g, p);
        uint32_t tout;
        const uint32_t cl = blk_carry<NT>(g, p, c, scan, ph, tout);
        c = tout;
        s = lane_post<SUB>(s, 0u, cl);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (k + j < len) r[k + j] = s[j];
    }
    __syncthreads();
}

// (lo64, cnt) column sums of A[0, la) B[0, lb) for columns [klo, klo + 4G),
// klo % 4 == 0, at plane index k - klo.  Work comes in tiles of 64
// consecutive column groups (256 columns) by up to kOneStripe rows of the
// shorter operand, which whole waves draw from an LDS counter: lane l of a
// tile owns group 64c + l and walks the stripe's row blocks that meet its
// rows in lock step with the other lanes, so the 8 row digits are one
// broadcast read and the lanes' windows of the other operand are consecutive
// 16-byte reads.  A lane adds its 4 column sums to the plane with LDS atomics
// after each tile.  Operands are read up to 12 words below 0 and 15 above
// their lengths (zero there).
template <int NT>
__device__ void blk_prod(const uint32_t* A, uint32_t la, const uint32_t* Bw, uint32_t lb, uint32_t klo, uint32_t G,
                         uint64_t* lo, uint32_t* cn, uint32_t* ctr) {
    constexpr uint32_t R = kOneRows;
    const uint32_t tid = threadIdx.x, lane = lane_id();
    for (uint32_t k = tid; k < 4 * G; k += NT) {
        lo[k] = 0;
        cn[k] = 0;
    }
    if (tid == 0) *ctr = 0;
    // rows from the shorter operand: a tile's lanes then share its row
    // blocks (a long row operand against a short one makes a thin diagonal
    // band, and a tile's lanes would walk 256 rows to use a few each)
    if (la > lb) {
        const uint32_t* t = A;
        A = Bw;
        Bw = t;
        const uint32_t tl = la;
        la = lb;
        lb = tl;
    }
    // stripe height: the tallest of 8 .. kOneStripe rows that still gives
    // every wave about three tiles (small products otherwise fall to one wave)
    const uint32_t C = (G + kWave - 1) / kWave;
    uint32_t H = kOneStripe;
    while (H > R && C * ((la + H - 1) / H) < 3 * (NT / kWave)) H >>= 1;
    const uint32_t NB = H / R, S = (la + H - 1) / H, T = C * S;
    __syncthreads();
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(ctr, 1u);
        t = readfirst(t);
        if (t >= T) break;
        const uint32_t st = t / C, c = t - st * C;
        const uint32_t r0 = st * H, g = kWave * c + lane, k0 = klo + 4 * g;
        // this lane's row blocks [jlo, jhi) of the stripe: rows lo_g <= i < hi_g
        // meet columns k0 .. k0+3
        uint32_t jlo = 0, jhi = 0;
        if (g < G) {
            const uint32_t lo_g = k0 + 1 > lb ? k0 + 1 - lb : 0u;
            const uint32_t hi_g = std::min(la, k0 + 4);
            if (hi_g > lo_g && hi_g > r0 && lo_g < r0 + H) {
                jlo = lo_g > r0 ? (lo_g - r0) / R : 0u;
                jhi = std::min(NB, (hi_g - r0 + R - 1) / R);
            }
        }
        const bool act = jhi > jlo;
        if (ballot(act) == 0) continue;
        const uint32_t jmin = NB - readfirst(wave_max_u32(act ? NB - jlo : 0u));
        const uint32_t jmax = readfirst(wave_max_u32(act ? jhi : 0u));
        uint64_t acc[4] = {0, 0, 0, 0};
        uint32_t cc[4] = {0, 0, 0, 0};
        const uint32_t* pa = A + r0 + R * jmin;
        const uint32_t* pw = Bw + ((int32_t)k0 - (int32_t)(r0 + R * jmin) - (int32_t)R);
        for (uint32_t j = jmin; j < jmax; j++) {
            if (j >= jlo && j < jhi) {
                const u32x4 a0 = *reinterpret_cast<const u32x4*>(pa);
                const u32x4 a1 = *reinterpret_cast<const u32x4*>(pa + 4);
                const u32x4 w0 = *reinterpret_cast<const u32x4*>(pw);
                const u32x4 w1 = *reinterpret_cast<const u32x4*>(pw + 4);
                const u32x4 w2 = *reinterpret_cast<const u32x4*>(pw + 8);
                // row r, column c reads window word R + c - r
                mac4(acc, cc, a0[0], w2[0], w2[1], w2[2], w2[3]);
                mac4(acc, cc, a0[1], w1[3], w2[0], w2[1], w2[2]);
                mac4(acc, cc, a0[2], w1[2], w1[3], w2[0], w2[1]);
                mac4(acc, cc, a0[3], w1[1], w1[2], w1[3], w2[0]);
                mac4(acc, cc, a1[0], w1[0], w1[1], w1[2], w1[3]);
                mac4(acc, cc, a1[1], w0[3], w1[0], w1[1], w1[2]);
                mac4(acc, cc, a1[2], w0[2], w0[3], w1[0], w1[1]);
                mac4(acc, cc, a1[3], w0[1], w0[2], w0[3], w1[0]);
            }
            pa += R;
            pw -= R;
        }
        if (act) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t idx = 4 * g + q;
                const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long*>(lo + idx),
                                               (unsigned long long)acc[q]);
                const uint32_t add = cc[q] + (old + acc[q] < old ? 1u : 0u);
                if (add) atomicAdd(cn + idx, add);
            }
        }
    }
    __syncthreads();
}

// Digits of V = sum_{klo <= k < khi} S_k B^k + X B^xoff (X: xlen digits at
// xs, none when xlen = 0) in [olo, olo + olen), the carry chain from klo with
// carry-in 0, stored at out[k - olo].  neg (klo == 0): the digits of
// (U - V) mod B^(olo + olen) instead (U: ulen digits at us, or none), taken
// as ~V + U + 1 -- a second carry chain whose wave aggregates, for either
// carry-in of the first chain, meet the first chain's in the same LDS round,
// so a tile costs one barrier either way.
template <int NT>
__device__ void blk_digits(const uint64_t* lo, const uint32_t* cn, uint32_t klo, uint32_t khi, const uint32_t* xs,
                           uint32_t xlen, uint32_t xoff, bool neg, const uint32_t* us, uint32_t ulen, uint32_t* out,
                           uint32_t olo, uint32_t olen, uint32_t* scan, uint32_t& ph, bool dec = false) {
    const uint32_t kend = olo + olen, w = blk_wave(), lane = lane_id();
    uint32_t c1 = 0, c2 = 1;   // carries into the tile: the sum's, the negation's +1
    for (uint32_t base = klo; base < kend; base += 4 * NT) {
        const uint32_t k = base + 4 * threadIdx.x;
        uint32_t* sl = scan + (ph & 1u) * kOneWaves;
        ph++;
        // a wave whose digits all lie past the window (the waves after it
        // too) publishes an empty aggregate and only meets the barrier
        if (base + 4 * kWave * w >= kend) {
            if (lane == 0) sl[w] = 0;
            __syncthreads();
            continue;
        }
        // columns k-4 .. k+3, and the added operand's digits at k-2 .. k+3
        uint32_t l32[8], h32[8], c32[8], x32[8], m32[8];
        const bool live = k < kend;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int64_t kk = (int64_t)k - 4 + j;
            l32[j] = h32[j] = c32[j] = x32[j] = 0;
            if (live && kk >= (int64_t)klo && kk < (int64_t)khi) {
                const uint64_t sv = lo[kk - klo];
                l32[j] = (uint32_t)sv;
                h32[j] = (uint32_t)(sv >> 32);
                c32[j] = cn[kk - klo];
            }
            if (live && j >= 2 && kk >= (int64_t)xoff && kk < (int64_t)xoff + xlen) x32[j] = xs[kk - xoff];
            // dec: B^kend - B^olo added, i.e. one subtracted at olo (mod B^kend)
            m32[j] = (dec && live && kk >= (int64_t)olo && kk < (int64_t)kend) ? 0xFFFFFFFFu : 0u;
        }
        // t for columns k-2 .. k+3 (< 5 2^32), y for k-1 .. k+3, z for k .. k+3
        uint32_t a[8], e[8], bb[8], f[8];
#pragma unroll
        for (int j = 2; j < 8; j++) {
            const uint64_t t = (uint64_t)l32[j] + h32[j - 1] + c32[j - 2] + x32[j] + m32[j];
            a[j] = (uint32_t)t;
            e[j] = (uint32_t)(t >> 32);
        }
#pragma unroll
        for (int j = 3; j < 8; j++) {
            const uint64_t y = (uint64_t)a[j] + e[j - 1];
            bb[j] = (uint32_t)y;
            f[j] = (uint32_t)(y >> 32);
        }
        u32x4 d;
        uint32_t cl = 0, all = 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t z = (uint64_t)bb[4 + j] + f[3 + j] + cl;
            d[j] = (uint32_t)z;
            cl = (uint32_t)(z >> 32);
            all &= d[j];
        }
        const bool g1 = cl != 0, p1 = all == 0xFFFFFFFFu;
        uint32_t wout;
        (void)wave_carry(g1, p1, 0u, wout);
        const uint32_t P1 = ballot(p1) == ~0ull ? 1u : 0u;
        u32x4 s2[2];
        bool g2[2], p2[2];
        uint32_t agg = wout | (P1 << 1);
        if (neg) {
            u32x4 uu;
#pragma unroll
            for (int j = 0; j < 4; j++) uu[j] = (live && k + j < ulen) ? us[k + j] : 0u;
#pragma unroll
            for (int cin = 0; cin < 2; cin++) {
                uint32_t dummy;
                const u32x4 dc = lane_post<false>(d, 0u, wave_carry(g1, p1, (uint32_t)cin, dummy));
                u32x4 nd;
#pragma unroll
                for (int j = 0; j < 4; j++) nd[j] = ~dc[j];
                lane_pre<false>(nd, uu, 0u, s2[cin], g2[cin], p2[cin]);
                uint32_t o2;
                (void)wave_carry(g2[cin], p2[cin], 0u, o2);
                const uint32_t A2 = ballot(p2[cin]) == ~0ull ? 1u : 0u;
                agg |= (o2 | (A2 << 1)) << (2 + 2 * cin);
            }
        }
        if (lane == 0) sl[w] = agg;
        __syncthreads();
        uint32_t cw1 = 0, cw2 = 0;
#pragma unroll
        for (int j = 0; j < NT / kWave; j++) {
            if (j == (int)w) {
                cw1 = c1;
                cw2 = c2;
            }
            const uint32_t sj = sl[j];
            const uint32_t sh = 2 + 2 * c1;   // the negation chain's aggregate for this wave's carry-in
            c2 = ((sj >> sh) & 1u) | ((sj >> (sh + 1)) & c2);
            c1 = (sj & 1u) | ((sj >> 1) & c1);
        }
        uint32_t dummy;
        d = lane_post<false>(d, 0u, wave_carry(g1, p1, cw1, dummy));
        if (neg) {
            const int ci = (int)cw1;
            d = lane_post<false>(ci ? s2[1] : s2[0], 0u, wave_carry(ci ? g2[1] : g2[0], ci ? p2[1] : p2[0], cw2, dummy));
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t kk = k + j;
            if (kk >= olo && kk < kend) out[kk - olo] = d[j];
        }
    }
    __syncthreads();
}

// X = floor(B^(2 p0) / (T + 1)) as p0 + 1 digits, T = the top p0 digits of V
// (zeros below them when NV < p0), 3 <= p0 <= 31 -- one wave.  With T >=
// B^(p0-1): X B^(NV-p0) <= B^(NV+p0) / V and B^(NV+p0) - V X < B^(NV+1) + V,
// the bound every Newton iterate keeps, so the steps start from X at p0.  Long
// division of B^(2 p0) by the normalised T + 1 (Knuth, TAOCP vol. 2, 4.3.1,
// algorithm D): remainder digit i in lane i (positions 0 .. 2 p0), one
// quotient digit a step from the top by the two-digit estimate and its test,
// the product subtracted as two borrow chains (its low words at the divisor's
// positions, its high words one position up), at most one add-back.  dn:
// p0 + 1 words of LDS for the normalised divisor.
__device__ void wave_recip_top(const uint32_t* V, uint32_t NV, uint32_t p0, uint32_t* dn, uint32_t* X) {
    const uint32_t lane = lane_id();
    uint32_t t = 0;
    if (lane < p0) {
        const int32_t k = (int32_t)NV - (int32_t)p0 + (int32_t)lane;
        t = k >= 0 ? V[k] : 0u;
    }
    // D = T + 1; T = B^p0 - 1 gives X = B^p0
    uint32_t cout;
    const uint32_t inc = wave_carry(false, lane < p0 && t == 0xFFFFFFFFu, 1u, cout);
    if (__builtin_amdgcn_readlane(inc, p0)) {
        if (lane <= p0) X[lane] = lane == p0 ? 1u : 0u;
        return;
    }
    const uint32_t dd = lane < p0 ? t + inc : 0u;
    const uint32_t s = (uint32_t)__builtin_clz(__builtin_amdgcn_readlane(dd, p0 - 1));   // top digit nonzero
    const uint32_t up = __shfl_up(dd, 1);
    const uint32_t below = lane ? up : 0u;
    const uint32_t dnv = s ? (dd << s) | (below >> (32 - s)) : dd;
    if (lane < p0) dn[lane + 1] = dnv;
    if (lane == 0) dn[0] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t vt = __builtin_amdgcn_readlane(dnv, p0 - 1), v2 = __builtin_amdgcn_readlane(dnv, p0 - 2);
    const uint32_t rv = (uint32_t)(0xFFFFFFFFFFFFFFFFull / vt - (1ull << 32));
    // the numerator B^(2 p0) 2^s: one digit, at position 2 p0 <= 62
    uint32_t r = lane == 2 * p0 ? 1u << s : 0u;
    for (int32_t j = (int32_t)p0; j >= 0; j--) {
        // divisor digit lane - j (the product's low words) and lane - j - 1
        // (its high words) at this lane
        const int32_t a = (int32_t)lane - j;
        const uint32_t da = (a >= 0 && a < (int32_t)p0) ? dn[a + 1] : 0u;
        const uint32_t db = (a >= 1 && a <= (int32_t)p0) ? dn[a] : 0u;
        const uint32_t top = __builtin_amdgcn_readlane(r, j + p0);
        const uint32_t nx = __builtin_amdgcn_readlane(r, j + p0 - 1);
        const uint32_t nx2 = __builtin_amdgcn_readlane(r, j + p0 - 2);
        // top <= vt: the remainder is below B times the divisor
        uint32_t qh;
        uint64_t rh;
        if (top >= vt) {
            qh = 0xFFFFFFFFu;
            rh = (uint64_t)nx + vt;
        } else {
            uint32_t q, rr;
            div21(top, nx, vt, rv, q, rr);
            qh = q;
            rh = rr;
        }
        for (int it = 0; it < 2; it++) {
            if (rh < (1ull << 32) && (uint64_t)qh * v2 > ((rh << 32) | nx2)) {
                qh--;
                rh += vt;
            }
        }
        const uint32_t lo = (uint32_t)((uint64_t)qh * da);
        const uint32_t hi = (uint32_t)(((uint64_t)qh * db) >> 32);
        uint32_t c1, c2;
        const uint32_t e1 = wave_carry(r < lo, r == lo, 0u, c1);
        const uint32_t y = r - lo - e1;
        const uint32_t e2 = wave_carry(y < hi, y == hi, 0u, c2);
        r = y - hi - e2;
        if (c1 | c2) {   // the estimate was one too large: add the divisor back
            qh--;
            const uint32_t sm = r + da;
            uint32_t c3;
            const uint32_t e3 = wave_carry(sm < r, sm == 0xFFFFFFFFu, 0u, c3);
            r = sm + e3;
        }
        if (lane == 0) X[j] = qh;
    }
}

template <int NT>
__global__ void __launch_bounds__(NT) k_div_one(const uint32_t* __restrict__ us, const uint32_t* __restrict__ vs,
                                                         uint32_t mu, uint32_t mv, uint32_t* __restrict__ qs,
                                                         uint32_t* __restrict__ rs, uint8_t* __restrict__ bad,
                                                         uint32_t lds_words) {
    extern __shared__ __attribute__((aligned(16))) uint32_t sm[];
    const uint32_t tid = threadIdx.x;
    const uint64_t inst = blockIdx.x;
    const uint32_t* u = us + inst * mu;
    const uint32_t* v = vs + inst * mv;
    uint32_t* q = qs + inst * mu;
    uint32_t* r = rs + inst * mv;
    for (uint32_t k = 4 * tid; k < lds_words; k += 4 * NT)
        *reinterpret_cast<u32x4*>(sm + k) = u32x4{0u, 0u, 0u, 0u};
    // significant lengths (reduced through the last 4 * kOneWaves words, past
    // every layout)
    uint32_t su = 0, sv = 0;
    for (uint32_t k = tid; k < mu; k += NT)
        if (u[k]) su = k + 1;
    for (uint32_t k = tid; k < mv; k += NT)
        if (v[k]) sv = k + 1;
    uint32_t* red0 = sm + lds_words - 4 * kOneWaves;
    su = wave_max_u32(su);
    sv = wave_max_u32(sv);
    __syncthreads();   // the zeroing is done
    if (lane_id() == 0) {
        red0[blk_wave()] = su;
        red0[kOneWaves + blk_wave()] = sv;
    }
    __syncthreads();
    su = sv = 0;
#pragma unroll
    for (int j = 0; j < NT / kWave; j++) {
        su = std::max(su, red0[j]);
        sv = std::max(sv, red0[kOneWaves + j]);
    }
    if (sv == 0) {   // zero divisor: q = r = 0
        for (uint32_t k = tid; k < mu; k += NT) q[k] = 0;
        for (uint32_t k = tid; k < mv; k += NT) r[k] = 0;
        if (tid == 0) bad[inst] = 1;
        return;
    }
    const uint32_t NV = std::max<uint32_t>(sv, 3), s = NV - sv, NU = su + s;
    const uint32_t P = (su + 1 > sv ? su + 1 - sv : 1u) + 1, PP = P + 2, XC = PP + 2;
    const OneLayout L = one_layout(NV, P, NU);
    uint32_t* V = sm + L.V;
    uint32_t* E = sm + L.E;
    uint32_t* Xb[2] = {sm + L.X0, sm + L.X1};
    uint32_t* U = sm + L.U;
    uint64_t* plo = reinterpret_cast<uint64_t*>(sm + L.lo);
    uint32_t* pcn = sm + L.cnt;
    uint32_t* ctr = sm + L.ctr;
    uint32_t* red = sm + L.red;
    uint32_t* scan = sm + L.scan;
    uint32_t ph = 0;
    for (uint32_t k = tid; k < sv; k += NT) V[s + k] = v[k];
    for (uint32_t k = tid; k < su; k += NT) U[s + k] = u[k];
    __syncthreads();
    // Long quotients come in two halves, q = q1 B^k + q0 (as in bi_div.hip):
    // q1 = floor(U2 / V), U2 = floor(U / B^k), at most P - k - 1 digits; r1
    // < V and U3 = r1 B^k + (U mod B^k) < V B^k give q0 = floor(U3 / V) <
    // B^k; both from one inverse at T = max(P1, P0) + 2 digits.
    const bool split = P >= (NT == 1024 ? kOneSplitMinP1024 : kOneSplitMinP256);
    const uint32_t k_lo = split ? (P - 1) / 2 : 0u;
    const uint32_t P1 = P - k_lo, P0 = k_lo + 2;
    const uint32_t T = split ? std::max(P1, P0) + 2 : PP;
    // the start: X at the schedule's largest precision p0 <= kOneDirect
    uint32_t p0 = T;
    while (p0 > kOneDirect) p0 = (p0 + 3) / 2;
    if (blk_wave() == 0) wave_recip_top(V, NV, p0, reinterpret_cast<uint32_t*>(plo), Xb[0]);
    __syncthreads();
    // Newton steps: the precisions of the chain T, (T + 3) / 2, ... above
    // p0, in rising order; the one after p is the element whose successor is p
    uint32_t p = p0;
    int cur = 0;
    while (p < T) {
        uint32_t p2 = T;
        while ((p2 + 3) / 2 > p) p2 = (p2 + 3) / 2;
        const uint32_t* X = Xb[cur];
        uint32_t* X2 = Xb[cur ^ 1];
        const uint32_t c = NV >= p + 2 ? NV - p - 2 : 0u;
        const uint32_t lenE = NV + 2 - c;
        const uint32_t d = NV + 2 * p - p2 - c;
        // E / B^c = digits [c, NV + 2) of (-V X) mod B^(NV+2).  V's digits
        // below cw = c - p - 1 move E / B^c by less than one unit (their
        // product with X is below B^c), so the product takes V's digits [cw,
        // NV) only, staged in E with zeros around them: E / B^c comes out at
        // most one unit high, X' at most one unit high, and the carry pass of
        // X' subtracts one, which keeps X a lower bound.  Only where the
        // window saves kOneWindowMin digit products: the staging is a pass
        const uint32_t cw = (c > p + 1 && (c - p - 1) * (p + 1) >= kOneWindowMin) ? c - p - 1 : 0u;
        {
            const uint32_t* Vh = V;
            const uint32_t nvh = NV - cw;
            if (cw) {
                for (uint32_t k = tid; k < nvh + 16; k += NT) E[k] = k < nvh ? V[cw + k] : 0u;
                __syncthreads();
                Vh = E;
            }
            const uint32_t khi = std::min(nvh + p, nvh + 2);
            blk_prod<NT>(Vh, nvh, X, p + 1, 0u, (khi + 3) / 4, plo, pcn, ctr);
            // the staged digits above E's lenE digits back to zero (E is the
            // next product's operand, read up to 15 words past its length);
            // the carry pass writes only below lenE
            if (cw)
                for (uint32_t k = lenE + tid; k < nvh + 16; k += NT) E[k] = 0u;
            blk_digits<NT>(plo, pcn, 0u, khi, nullptr, 0u, 0u, true, nullptr, 0u, E, c - cw, lenE, scan, ph);
        }
        // X' = X B^(p2 - p) + digits [d, d + p2 + 2) of X (E / B^c) from the
        // top columns: the shifted X enters the carry pass at column
        // d + p2 - p, the window stops at XC digits (X' < B^(p2+1))
        {
            const uint32_t ktop = p + lenE, klo = (d - 2) & ~3u;   // d >= min(p + 4, NV + 2) >= 5
            blk_prod<NT>(X, p + 1, E, lenE, klo, (ktop - klo + 3) / 4, plo, pcn, ctr);
            blk_digits<NT>(plo, pcn, klo, ktop, X, p + 1, d + p2 - p, false, nullptr, 0u, X2, d, std::min(p2 + 2, XC),
                       scan, ph, cw != 0);
        }
        cur ^= 1;
        p = p2;
    }
    uint32_t* Wb = Xb[cur];   // W: T + 1 digits
    uint32_t* Q = Xb[cur ^ 1];
    uint32_t* Rr = E;
    // Half 0 (two halves only): q1 from U2 = U / B^k (U's digits from k up)
    // into Q from digit k up.  Half 1: the quotient of U (one piece) into Q,
    // or q0 from U3 (r1 B^k + (U mod B^k), written over U from digit k up)
    // into W's buffer, whose digits the product no longer needs once it has
    // run.  One loop, so the passes' code is inlined once.
    for (uint32_t h = split ? 0u : 1u; h < 2; h++) {
        const uint32_t off = h ? 0u : k_lo;
        const uint32_t nu = h ? (split ? NV + 1 + k_lo : NU) : NU - k_lo;
        const uint32_t pq = h ? (split ? P0 : P) : P1;   // pq - 1: the most quotient digits
        const uint32_t pw = h ? (split ? k_lo : P) : P1; // quotient digits kept
        const uint32_t* Un = U + off;
        uint32_t* Qb = (h && split) ? Wb : Q;
        uint32_t* Qd = Qb + off;
        // the quotient: digits [NV + T, NV + T + pw) of Un W, top columns only
        {
            const uint32_t LU = std::max<uint32_t>(1, std::min(nu, pq + NV - 2));
            const uint32_t ktop = LU + T, klo = (NV + T - 2) & ~3u;
            if (ktop > klo) blk_prod<NT>(Un, LU, Wb, T + 1, klo, (ktop - klo + 3) / 4, plo, pcn, ctr);
            blk_digits<NT>(plo, pcn, klo, ktop > klo ? ktop : klo, nullptr, 0u, 0u, false, nullptr, 0u, Qd, NV + T, pw,
                           scan, ph);
        }
        // every other digit of the quotient's buffer cleared: it is a product
        // operand next
        for (uint32_t k = tid; k < XC + 4; k += NT)
            if (k < off || k >= off + pw) Qb[k] = 0;
        __syncthreads();
        // Rr = (Un - Qd V) mod B^(NV+1) (the low product's negation pass adds
        // Un), then the delta fix, twice: Rr >= V -> Qd += 1, Rr -= V
        {
            const uint32_t khi = std::min(pw + NV - 1, NV + 1);
            blk_prod<NT>(Qd, pw, V, NV, 0u, (khi + 3) / 4, plo, pcn, ctr);
            blk_digits<NT>(plo, pcn, 0u, khi, nullptr, 0u, 0u, true, Un, std::min(nu, NV + 1), Rr, 0u, NV + 1, scan,
                           ph);
        }
        for (int it = 0; it < 2; it++) {
            uint32_t hit = 0;
            for (uint32_t k = tid; k < NV + 1; k += NT)
                if (Rr[k] != V[k]) hit = k + 1;
            hit = blk_max<NT>(hit, red, ph);
            const bool ge = hit == 0 || Rr[hit - 1] > V[hit - 1];
            __syncthreads();
            if (!ge) break;
            blk_addsub<false, NT>(Qd, pw, 0u, nullptr, 0u, 1u, Qd, pw, scan, ph);
            blk_addsub<true, NT>(Rr, NV + 1, 0u, V, NV, 0u, Rr, NV + 1, scan, ph);
        }
        if (h == 0) {   // U3 = r1 B^k + (U mod B^k)
            for (uint32_t k = tid; k < NU - k_lo; k += NT) U[k_lo + k] = k < NV + 1 ? Rr[k] : 0u;
            __syncthreads();
        }
    }
    // q = q1 B^k + q0 (digits from P up are zero), r = R / B^s
    for (uint32_t k = tid; k < mu; k += NT) q[k] = k < k_lo ? Wb[k] : (k < P ? Q[k] : 0u);
    for (uint32_t k = tid; k < mv; k += NT) r[k] = k + s < NV + 1 ? Rr[k + s] : 0u;
    if (tid == 0) bad[inst] = 0;
}

// Launches the one-workgroup division of n instances on st with NT lanes a
// workgroup (256 or 1024); bad[i] = 1 marks a zero divisor (q = r = 0
// there).  No host synchronisation.
hipError_t divone_launch_nt(int nt, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs,
                            uint32_t* qs, uint32_t* rs, uint8_t* bad, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (!divone_fits(mu, mv) || n > 0x7FFFFFFFull || (nt != 256 && nt != 1024)) return hipErrorInvalidValue;
    static const bool attr = [] {
        set_max_lds(reinterpret_cast<const void*>(&k_div_one<256>), (int)kOneLdsMax);
        set_max_lds(reinterpret_cast<const void*>(&k_div_one<1024>), (int)kOneLdsMax);
        return true;
    }();
    (void)attr;
    const uint32_t words = one_lds_words(mu, mv) + 4 * kOneWaves;
    if (nt == 256)
        hipLaunchKernelGGL(k_div_one<256>, dim3((uint32_t)n), dim3(256), (size_t)words * 4, st, us, vs, mu, mv, qs, rs,
                           bad, words);
    else
        hipLaunchKernelGGL(k_div_one<1024>, dim3((uint32_t)n), dim3(1024), (size_t)words * 4, st, us, vs, mu, mv, qs,
                           rs, bad, words);
    return hipGetLastError();
}

// Workgroup size: 1024 lanes keep a long division's products busy when
// there is about one division a CU; 256 lanes when operands are short or
// there are more divisions than CUs (four workgroups share a CU, each
// barrier waits for four waves, not sixteen).  Kernel time on one MI355X,
// 256 / 1024 lanes (us):
//   digits        n = 16        256          1024          4096
//   64 / 32       23 / 25      22 / 25      27 / 93       79 / 362
//   256 / 128     45 / 39      45 / 40      58 / 152     178 / 597
//   512 / 256     67 / 56      68 / 58      96 / 222     304 / 867
//   2048 / 1024  189 / 132    198 / 143    438 / 555    1435 / 2120
//   4096 / 2048  474 / 273    498 / 304   1929 / 1168   7623 / 4384
// (measured before long quotients came in two halves, which took the
// 1024-lane 2048 / 1024 and 4096 / 2048 calls at n = 16 from 139 and 262 us
// to 127 and 218 us, and 1024 divisions of 2048 / 1024 digits on 256 lanes
// from 402 to 335 us: both widths gain, the choice stands)
hipError_t divone_launch(uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs, uint32_t* qs,
                         uint32_t* rs, uint8_t* bad, hipStream_t st) {
    const uint32_t len = std::max(mu, mv);
    const bool small = len <= kOneSmallDigits || (n > 256 && len <= 2048);
    return divone_launch_nt(small ? 256 : 1024, n, mu, mv, us, vs, qs, rs, bad, st);
}

// Whether a batch takes the one-launch division: up to n_max instances whose
// layout fits one workgroup, and for operands past 2048 digits at most 2048
// of them (from there the batched pipeline's transform products catch up:
// 4096 divisions of 3000 / 1500 digits 2.61 ms in one launch, 2.47 ms
// batched; 6000 / 3000 digits 7.03 and 6.44 ms; 4096 / 2048 digits 4.37 and
// 4.50 ms.  Up to 2048 digits the one launch stays ahead: 8192 divisions of
// 2048 / 1024 digits 2.75 ms, batched 2.86 ms).
bool divone_wanted(uint64_t n, uint32_t mu, uint32_t mv, uint32_t n_max) {
    return n <= n_max && divone_fits(mu, mv) && (n <= 2048 || std::max(mu, mv) <= 2048);
}

}  // namespace bi
