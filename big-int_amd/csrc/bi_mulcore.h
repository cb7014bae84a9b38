// bi_mulcore.h -- the per-lane convolution shared by the uniform and the
// ragged multiply kernels.
//
// A team of lanes multiplies one pair of integers staged in LDS (the row
// operand A and the zero-padded window operand, layouts below) and leaves
// each lane's 96-bit column sums in registers, then folds them into the
// team's column arrays.  Lane schedules (which 4-column groups a lane owns and
// which R-row blocks it walks) are computed by sched_low_r / sched_rect_r.
#pragma once
#include "bi_device.h"

namespace bi {

struct LaneSched {
    uint32_t stepsA, stepsB;  // 4x4 blocks walked in the first / second column group
    uint32_t kA, kB;          // first column of each group
    uint32_t tB0;             // first row block of the second group
    uint32_t tA0;             // first row block of the first group
};

__device__ __forceinline__ void store_cols(uint32_t* CL, uint32_t* CH, uint32_t* CC, uint32_t k, const uint64_t* acc,
                                           const uint32_t* cnt) {
    *reinterpret_cast<u32x4*>(CL + k) = u32x4{(uint32_t)acc[0], (uint32_t)acc[1], (uint32_t)acc[2], (uint32_t)acc[3]};
    *reinterpret_cast<u32x4*>(CH + k) = u32x4{(uint32_t)(acc[0] >> 32), (uint32_t)(acc[1] >> 32),
                                              (uint32_t)(acc[2] >> 32), (uint32_t)(acc[3] >> 32)};
    *reinterpret_cast<u32x4*>(CC + k) = u32x4{cnt[0], cnt[1], cnt[2], cnt[3]};
}

// ---------------------------------------------------------------------------
// R-row blocks: a step multiplies R digits of the row operand (rows R*t ..
// R*t+R-1) by a sliding (R+4)-digit window of the window operand into the
// lane's 4 columns: 4R multiply-adds per (R/4 + R/4 + 1) ds_read_b128.
// Window operand layout: b_j at B[R + j], B[0..R-1] = 0 and zeros above.
// Schedules (4-column groups g, row blocks t; NA, NB in 4-digit units with
// 4*NA and 4*NB multiples of R):
//   low : lane l owns groups l and G-1-l (rows <= column), +-1 block balance;
//   rect: job j < NB pairs the rising group j with the falling group NA+j,
//         NB <= j < NA takes the plateau group j; every job walks the same
//         number of blocks (4*NB/R + 1).
template <int R>
__device__ __forceinline__ LaneSched sched_low_r(uint32_t l, uint32_t G) {
    LaneSched s = {0, 0, 0, 0, 0, 0};
    if (l < (G + 1) / 2) {
        const uint32_t gB = G - 1 - l;
        s.stepsA = (4 * l + 3) / R + 1;
        s.kA = 4 * l;
        s.stepsB = (gB == l) ? 0u : (4 * gB + 3) / R + 1;
        s.kB = 4 * gB;
    }
    return s;
}

// first row block of column group g whose window still reaches b (< 4*NB)
template <int R>
__device__ __forceinline__ uint32_t rect_tlo(uint32_t g, uint32_t NB) {
    if (g < NB) return 0;
    // floor((4(g-NB) - R + 1) / R) + 1, non-negative for g >= NB
    return (4 * (g - NB) + 1) / R;
}

template <int R>
__device__ __forceinline__ LaneSched sched_rect_r(uint32_t j, uint32_t NA, uint32_t NB) {
    LaneSched s = {0, 0, 0, 0, 0, 0};
    const uint32_t nblk = 4 * NA / R;   // row blocks
    if (j < NB) {
        s.stepsA = (4 * j + 3) / R + 1;
        s.kA = 4 * j;
        const uint32_t g2 = NA + j;
        const uint32_t t0 = rect_tlo<R>(g2, NB);
        s.stepsB = nblk - t0;
        s.kB = 4 * g2;
        s.tB0 = t0;
    } else if (j < NA) {
        const uint32_t t0 = rect_tlo<R>(j, NB);
        const uint32_t t1c = (4 * j + 3) / R;
        const uint32_t t1 = t1c < nblk - 1 ? t1c : nblk - 1;
        s.stepsA = t1 - t0 + 1;
        s.kA = 4 * j;
        s.tA0 = t0;
    }
    return s;
}

// ---------------------------------------------------------------------------
// Team LDS layout of the uniform and ragged multiply kernels (words; M8 =
// digits rounded up to 8, KW = number of product columns kept: M8 for the low
// product, 2*M8 for the full one):
//   operands   A = U:          a_i at A[i], zeros up to M8+12;
//              W = U + M8+12:  W[0..20) = 0, b_j at W[20+j], zeros up to
//                              b index M8+16 (2*M8+48 words in all);
//   columns    E = U, O = U + CW (CW = KW+8), written once every lane is done
//              with the operands.  The lane owning the 4-column group at
//              column k folds its four 96-bit column sums into one 6-digit
//              number (sum of col_{k+j} * B^j) and stores it at k: groups
//              with k % 8 == 0 in E, the others in O, 8 words each (6 digits
//              and 2 zeros), so neither array overlaps itself.  Digit k of the
//              product before carries is then E[k] + O[k] < 2^33.
__host__ __device__ inline uint32_t team_col_stride(uint32_t KW) { return KW + 8; }
__host__ __device__ inline uint32_t team_words(uint32_t M8, uint32_t KW) {
    const uint32_t ops = 2 * M8 + 48, cols = 2 * team_col_stride(KW);
    const uint32_t u = ops > cols ? ops : cols;
    return 4 + ((u + 3) & ~3u);
}
constexpr uint32_t kWinPad = 20;   // zero words in front of b_0 in W

// Walk of one lane over its blocks with running LDS pointers.  A step reads R
// digits of a at A[pa] and the (R+4)-digit window Wr[pb .. pb+R+3] (Wr = W +
// 16 - 2R, so window position pb = y + R + 4 holds b_{y-R} .. b_{y+3} for the
// block whose row digits start R*t below column kb, y = kb - R*t); then pa
// moves up and pb down by R.  pb saturates at 0, where the window is all zero
// padding: a lane past the end of its blocks (or with none) only adds zero
// products, so the loop needs no per-lane guard.  At the switch step the
// first group's sums are recorded in `snap` (12 words of global scratch owned
// by the lane: four 64-bit sums, then four overflow counts) and the walk
// jumps to the second group.
template <int R>
__device__ __forceinline__ void conv_walk(const uint32_t* A, const uint32_t* Wr, const LaneSched sc, uint32_t S,
                                          uint32_t* snap, uint64_t* acc, uint32_t* cnt) {
    const uint32_t total = sc.stepsA + sc.stepsB;
    uint32_t pa = 0, pb = 0, pa2 = 0, pb2 = 0, sw = 0xFFFFFFFFu;
    if (total) {
        pa = R * sc.tA0;
        pb = sc.kA - R * sc.tA0 + R + 4;
    }
    if (sc.stepsB) {
        sw = sc.stepsA;
        pa2 = R * sc.tB0;
        pb2 = sc.kB - R * sc.tB0 + R + 4;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        acc[c] = 0;
        cnt[c] = 0;
    }
    for (uint32_t s = 0; s < S; s++) {
        uint32_t av[R], wv[R + 4];
#pragma unroll
        for (int q = 0; q < R / 4; q++) {
            const u32x4 x = *reinterpret_cast<const u32x4*>(A + pa + 4 * q);
            av[4 * q] = x[0];
            av[4 * q + 1] = x[1];
            av[4 * q + 2] = x[2];
            av[4 * q + 3] = x[3];
        }
#pragma unroll
        for (int q = 0; q < R / 4 + 1; q++) {
            const u32x4 x = *reinterpret_cast<const u32x4*>(Wr + pb + 4 * q);
            wv[4 * q] = x[0];
            wv[4 * q + 1] = x[1];
            wv[4 * q + 2] = x[2];
            wv[4 * q + 3] = x[3];
        }
#pragma unroll
        for (int r = 0; r < R; r++) mac4(acc, cnt, av[r], wv[R - r], wv[R + 1 - r], wv[R + 2 - r], wv[R + 3 - r]);
        pa += R;
        pb = __builtin_elementwise_sub_sat(pb, (uint32_t)R);
        if (s + 1 == sw) {
            uint64_t* s64 = reinterpret_cast<uint64_t*>(snap);
            s64[0] = acc[0];
            s64[1] = acc[1];
            s64[2] = acc[2];
            s64[3] = acc[3];
            *reinterpret_cast<u32x4*>(snap + 8) = u32x4{cnt[0], cnt[1], cnt[2], cnt[3]};
            pa = pa2;
            pb = pb2;
        }
    }
}

// The part of a lane schedule in row blocks [T0, T1), for a product whose
// rows are split over several workgroups (each adds its rows' products into
// planes of its own).  A group left without rows still owns its columns in
// this part's planes: their (zero) sums go to kz[0..nz).  A group B left
// alone moves into slot A.
__device__ __forceinline__ LaneSched clip_rows(LaneSched sc, uint32_t T0, uint32_t T1, uint32_t* kz, uint32_t& nz) {
    nz = 0;
    const bool hadA = sc.stepsA > 0, hadB = sc.stepsB > 0;
    uint32_t a = sc.tA0 > T0 ? sc.tA0 : T0, b = sc.tA0 + sc.stepsA < T1 ? sc.tA0 + sc.stepsA : T1;
    sc.stepsA = b > a ? b - a : 0u;
    sc.tA0 = a;
    a = sc.tB0 > T0 ? sc.tB0 : T0;
    b = sc.tB0 + sc.stepsB < T1 ? sc.tB0 + sc.stepsB : T1;
    sc.stepsB = b > a ? b - a : 0u;
    sc.tB0 = a;
    if (hadA && sc.stepsA == 0) kz[nz++] = sc.kA;
    if (hadB && sc.stepsB == 0) kz[nz++] = sc.kB;
    if (sc.stepsA == 0 && sc.stepsB) {
        sc.kA = sc.kB;
        sc.tA0 = sc.tB0;
        sc.stepsA = sc.stepsB;
        sc.stepsB = 0;
    }
    return sc;
}

// conv_walk with the first group's sums recorded in their own slots of global
// column planes (CL/CH/CC at kA) instead of a per-lane snapshot; on return
// both groups' sums are in the planes (natural, unshifted layout).  CLIP (a
// schedule cut by clip_rows): the walk does not end where the window runs off
// the operand, so after the lane's last block the window is set to the zero
// padding explicitly.
template <int R, bool CLIP = false>
__device__ __forceinline__ void conv_walk_planes(const uint32_t* A, const uint32_t* Wr, const LaneSched sc, uint32_t S,
                                                 uint32_t* CL, uint32_t* CH, uint32_t* CC) {
    const uint32_t total = sc.stepsA + sc.stepsB;
    uint32_t pa = 0, pb = 0, pa2 = 0, pb2 = 0, sw = 0xFFFFFFFFu;
    if (total) {
        pa = R * sc.tA0;
        pb = sc.kA - R * sc.tA0 + R + 4;
    }
    if (sc.stepsB) {
        sw = sc.stepsA;
        pa2 = R * sc.tB0;
        pb2 = sc.kB - R * sc.tB0 + R + 4;
    }
    uint64_t acc[4] = {0, 0, 0, 0};
    uint32_t cnt[4] = {0, 0, 0, 0};
    for (uint32_t s = 0; s < S; s++) {
        uint32_t av[R], wv[R + 4];
#pragma unroll
        for (int q = 0; q < R / 4; q++) {
            const u32x4 x = *reinterpret_cast<const u32x4*>(A + pa + 4 * q);
            av[4 * q] = x[0];
            av[4 * q + 1] = x[1];
            av[4 * q + 2] = x[2];
            av[4 * q + 3] = x[3];
        }
#pragma unroll
        for (int q = 0; q < R / 4 + 1; q++) {
            const u32x4 x = *reinterpret_cast<const u32x4*>(Wr + pb + 4 * q);
            wv[4 * q] = x[0];
            wv[4 * q + 1] = x[1];
            wv[4 * q + 2] = x[2];
            wv[4 * q + 3] = x[3];
        }
#pragma unroll
        for (int r = 0; r < R; r++) mac4(acc, cnt, av[r], wv[R - r], wv[R + 1 - r], wv[R + 2 - r], wv[R + 3 - r]);
        pa += R;
        pb = __builtin_elementwise_sub_sat(pb, (uint32_t)R);
        if (s + 1 == sw) {
            store_cols(CL, CH, CC, sc.kA, acc, cnt);
            pa = pa2;
            pb = pb2;
        }
        if (CLIP && s + 1 == total) pb = 0;
    }
    if (total) {
        uint32_t kw = sc.kA;
        if (sc.stepsB) {
            const u32x4 l = *reinterpret_cast<const u32x4*>(CL + sc.kA);
            const u32x4 h = *reinterpret_cast<const u32x4*>(CH + sc.kA);
            const u32x4 c = *reinterpret_cast<const u32x4*>(CC + sc.kA);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint64_t rec = ((uint64_t)h[j] << 32) | l[j];
                const uint32_t br = acc[j] < rec;
                acc[j] -= rec;
                cnt[j] = cnt[j] - c[j] - br;
            }
            kw = sc.kB;
        }
        store_cols(CL, CH, CC, kw, acc, cnt);
    }
}

// The four column sums of the group at column k (k % 4 == 0) folded into 6
// digits, stored with 2 zero digits at E[k..k+8) (k % 8 == 0) or O[k..k+8).
__device__ __forceinline__ void put_group(uint32_t* E, uint32_t* O, uint32_t k, const uint64_t* acc,
                                          const uint32_t* cnt) {
    const uint32_t l0 = (uint32_t)acc[0], l1 = (uint32_t)acc[1], l2 = (uint32_t)acc[2], l3 = (uint32_t)acc[3];
    const uint32_t h0 = (uint32_t)(acc[0] >> 32), h1 = (uint32_t)(acc[1] >> 32), h2 = (uint32_t)(acc[2] >> 32),
                   h3 = (uint32_t)(acc[3] >> 32);
    uint32_t k1, ka, kb, kc;
    const uint32_t d1 = __builtin_addc(h0, l1, 0u, &k1);
    uint32_t t = __builtin_addc(cnt[0], h1, 0u, &ka);
    const uint32_t d2 = __builtin_addc(t, l2, k1, &kb);
    const uint32_t k2 = ka + kb;
    t = __builtin_addc(cnt[1], h2, 0u, &ka);
    t = __builtin_addc(t, l3, 0u, &kb);
    const uint32_t d3 = __builtin_addc(t, k2, 0u, &kc);
    const uint32_t k3 = ka + kb + kc;
    t = __builtin_addc(cnt[2], h3, 0u, &ka);
    const uint32_t d4 = __builtin_addc(t, k3, 0u, &kb);
    const uint32_t d5 = cnt[3] + ka + kb;
    uint32_t* dst = ((k & 7u) ? O : E) + k;
    *reinterpret_cast<u32x4*>(dst) = u32x4{l0, d1, d2, d3};
    *reinterpret_cast<u32x4*>(dst + 4) = u32x4{d4, d5, 0u, 0u};
}

// Writes a lane's groups after conv_walk (second group = running sums minus
// the recorded first-group sums).
__device__ __forceinline__ void put_lane_groups(uint32_t* E, uint32_t* O, const LaneSched sc, const uint32_t* snap,
                                                uint64_t* acc, uint32_t* cnt) {
    if (sc.stepsA + sc.stepsB == 0) return;
    if (sc.stepsB) {
        const uint64_t* s64 = reinterpret_cast<const uint64_t*>(snap);
        uint64_t ra[4];
        uint32_t rc[4];
        const u32x4 c4 = *reinterpret_cast<const u32x4*>(snap + 8);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ra[j] = s64[j];
            rc[j] = c4[j];
            const uint32_t br = acc[j] < ra[j];
            acc[j] -= ra[j];
            cnt[j] = cnt[j] - rc[j] - br;
        }
        put_group(E, O, sc.kA, ra, rc);
        put_group(E, O, sc.kB, acc, cnt);
    } else {
        put_group(E, O, sc.kA, acc, cnt);
    }
}

// Zero-fills what no group writes: E[K..lim), O[K+4..lim) (groups cover
// columns < K: E up to K, O from 4 up to K+4; K % 8 == 0) and O[0..4).
__device__ __forceinline__ void zero_groups(uint32_t* E, uint32_t* O, uint32_t K, uint32_t lim, uint32_t tl,
                                            uint32_t T) {
    for (uint32_t i = K + tl; i < lim; i += T) {
        E[i] = 0;
        if (i >= K + 4) O[i] = 0;
    }
    if (tl < 4) O[tl] = 0;
    if (tl == 0 && T < 4)
        for (uint32_t i = T; i < 4; i++) O[i] = 0;
}

// t = l + h + c (each < 2^32) as its low word and its high part (0..2)
__device__ __forceinline__ void tsplit(uint32_t l, uint32_t h, uint32_t c, uint32_t& lo, uint32_t& hi) {
    uint32_t c1, c2;
    const uint32_t s1 = __builtin_addc(l, h, 0u, &c1);
    lo = __builtin_addc(s1, c, 0u, &c2);
    hi = c1 + c2;
}

// Lane part of the recombination add over 4 digits: x[j] the low words,
// y[j] the high part of the digit below (0 at an instance start), fl bit j
// set when digit j starts an instance (the carry is cut there).  s = x + y
// + carry per digit; g / p: the lane generates / propagates a carry.
__device__ __forceinline__ void rec_lane(const u32x4 x, const u32x4 y, uint32_t fl, u32x4& s, bool& g, bool& p) {
    uint32_t c = 0, all = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        c &= ~(fl >> j) & 1u;
        s[j] = __builtin_addc(x[j], y[j], c, &c);
        all &= s[j];
    }
    g = c != 0;
    p = (fl == 0) && all == 0xFFFFFFFFu;
}

// Final digits: s + the carry e into the lane, cut at instance starts.
__device__ __forceinline__ u32x4 rec_post(const u32x4 s, uint32_t fl, uint32_t e) {
    u32x4 r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        e &= ~(fl >> j) & 1u;
        r[j] = s[j] + e;
        e &= (uint32_t)(s[j] == 0xFFFFFFFFu);
    }
    return r;
}

// Low words, high parts and start flags of the 4 digits at flat position qb
// (< ndig) of a run of consecutive instances of m digits whose teams' column
// arrays start at `base0` (E, with O = E + CW), one team every `words` words.
// VEC: m % 4 == 0, so the 4 digits share one instance and are read as 16-byte
// vectors.
template <bool VEC>
__device__ __forceinline__ void rec_digits(const uint32_t* base0, uint32_t words, uint32_t CW, uint32_t m,
                                           uint64_t magic, uint32_t qb, uint32_t ndig, u32x4& X, u32x4& Y,
                                           uint32_t& fl) {
    const uint32_t tm = (uint32_t)(((uint64_t)qb * magic) >> 32);
    uint32_t k = qb - tm * m;
    const uint32_t* base = base0 + (size_t)tm * words;
    fl = 0;
    if (VEC) {
        const u32x4 e4 = *reinterpret_cast<const u32x4*>(base + k);
        const u32x4 o4 = *reinterpret_cast<const u32x4*>(base + CW + k);
        const uint32_t km = k ? k - 1 : 0;
        uint32_t hp;
        (void)__builtin_addc(base[km], base[CW + km], 0u, &hp);
        uint32_t lo[4], hi[4];
#pragma unroll
        for (int j = 0; j < 4; j++) lo[j] = __builtin_addc(e4[j], o4[j], 0u, &hi[j]);
        X = u32x4{lo[0], lo[1], lo[2], lo[3]};
        Y = u32x4{k ? hp : 0u, hi[0], hi[1], hi[2]};
        fl = k ? 0u : 1u;
    } else {
        uint32_t hp = 0;
        if (k) (void)__builtin_addc(base[k - 1], base[CW + k - 1], 0u, &hp);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            X[j] = 0;
            Y[j] = 0;
            if (qb + j < ndig) {
                uint32_t hj;
                X[j] = __builtin_addc(base[k], base[CW + k], 0u, &hj);
                Y[j] = k ? hp : 0u;
                if (!k) fl |= 1u << j;
                hp = hj;
            }
            if (++k == m) {
                k = 0;
                base += words;
            }
        }
    }
}

}  // namespace bi
