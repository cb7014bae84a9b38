// bi_mul.hip -- batched multiply: rs = low m digits of as*bs per instance.
//
// Reference semantics: cuda/ker-mul.cu.h:147-666 (convMul1..5) compute the
// convolutions k = 0..m-1 only (the low m digits); the GMP check truncates the
// product to m words (cuda/helpers.h:161-187).
//
// Schedule.  The reference balances work by giving a thread the columns k and
// m-1-k (k+1 and m-k terms).  Here a lane owns two groups of FOUR consecutive
// columns and walks them in R x 4 blocks (R = 8 digits of a against a sliding
// 12-digit window of b, 5 ds_read_b128 for 32 multiply-adds):
//
//   truncated (low m of an m x m product): lane l owns column groups l and
//     G-1-l (G = M8/4), about G*4/R + 1 blocks in total;
//   full (both operands fit in N <= M8/2 digits, e.g. top halves zero): job j
//     owns the rising group j and the falling group N/4+j, N/R + 1 blocks
//     each -- only the N^2 useful products.
//
// The number of significant digits of each operand is found on the device, so
// an instance whose operands fit in half the width takes the full schedule and
// never multiplies the zero top halves.
//
// A lane walks its first group, records the column sums reached at the switch
// (in a per-lane global scratch slot, so that the column sums may later reuse
// the operands' LDS), keeps accumulating the second group in the same
// registers, and subtracts the record at the end.  A column sum is a 64-bit
// accumulator plus an overflow count (v_mad_u64_u32 carry-out, see bi::mac4),
// i.e. the reference's low/high/carry triple (cuda/ker-mul.cu.h:16-67).  The
// reference's l + (h<<1) + (c<<2) recombination (cuda/ker-mul.cu.h:197-199)
// becomes: each lane folds the four column sums of a group into one 6-digit
// number, even and odd groups go to two arrays (bi_mulcore.h), and ONE
// carry-look-ahead add per 256-digit tile adds the two arrays.
#include "bi_kernels.h"
#include "bi_mulcore.h"

#include <algorithm>

namespace bi {

struct MulGeom {
    uint32_t m, M8, G, T;          // digits, digits padded to 8, column groups, lanes per team
    uint32_t teams_per_block;      // instances per workgroup
    uint32_t tpw;                  // teams per wave (T <= 64) or 0 (team spans waves)
    uint32_t words;                // LDS words per team
    uint32_t block;                // threads per workgroup
    uint32_t CW;                   // column array stride (words)
    uint64_t magic;                // floor(2^32 / m) + 1: q / m = (q * magic) >> 32 for q < 2^18
};

// Team LDS layout: META[4] | U with the operand / column layout of
// bi_mulcore.h (team_words).  The first-group sums recorded at a lane's switch
// go to global scratch, so the two uses of U never overlap in time.
static MulGeom mul_geom(uint32_t m, uint32_t wpb) {
    MulGeom g;
    g.m = m;
    g.M8 = (m + 7) & ~7u;
    g.G = g.M8 / 4;
    g.T = g.M8 / 8;
    g.CW = team_col_stride(g.M8);
    g.words = team_words(g.M8, g.M8);
    g.magic = (1ull << 32) / m + 1;
    if (g.T <= 64) {
        g.tpw = 64 / g.T;
        g.teams_per_block = wpb * g.tpw;
        g.block = wpb * 64;
    } else {
        g.tpw = 0;
        g.teams_per_block = 1;
        g.block = ((g.T + 63) / 64) * 64;
    }
    return g;
}

// LDS ordering among the lanes of one team: a wave-local fence when a team
// (and every team sharing its LDS) lives in one wave, else the block barrier.
template <bool WAVE>
__device__ __forceinline__ void team_sync() {
    if (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Persistent.  WAVE (teams of <= 64 lanes): every wave runs on its own, unit
// u = tpw consecutive instances, wave w of block b takes units b*wpb + w,
// + gridDim.x*wpb, ... with its own LDS; no block barriers.  Otherwise one
// team spans the block and unit u = one instance, units b, b + gridDim.x, ...
// At least six waves a SIMD (at most 85 registers a lane, a few bytes
// spilled outside the walk): 131072 x 512 digits 0.710 -> 0.698 ms against
// the 86 registers (five waves) the compiler chose by itself; at eight waves
// (64 registers, 72 bytes spilled) 0.716 ms.
template <bool VEC, int R, bool WAVE>
__global__ void __launch_bounds__(1024, 6) k_mul(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                              uint32_t* __restrict__ rs, uint64_t n, MulGeom geo,
                                              uint32_t* __restrict__ snaps, uint64_t nunits) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t m = geo.m, M8 = geo.M8, T = geo.T, CW = geo.CW;
    const uint32_t wv = tid / 64, ln = tid % 64;
    uint32_t team, tl, per;
    bool member;
    uint32_t* tlds;
    uint64_t u0, ustride;
    if (WAVE) {
        team = ln / T;
        tl = ln % T;
        member = ln < geo.tpw * T;
        per = geo.tpw;
        tlds = lds + (size_t)wv * geo.tpw * geo.words;
        u0 = (uint64_t)blockIdx.x * (blockDim.x / 64) + wv;
        ustride = (uint64_t)gridDim.x * (blockDim.x / 64);
    } else {
        team = 0;
        tl = tid;
        member = tid < T;
        per = 1;
        tlds = lds;
        u0 = blockIdx.x;
        ustride = gridDim.x;
    }
    uint32_t* META = tlds + (size_t)team * geo.words;   // [0] prec a, [1] prec b
    uint32_t* U = META + 4;
    uint32_t* A = U;
    uint32_t* W = U + M8 + 12;
    uint32_t* E = U;
    uint32_t* O = U + CW;
    uint32_t* snap = snaps + ((size_t)blockIdx.x * blockDim.x + tid) * 12;

    for (uint64_t unit = u0; unit < nunits; unit += ustride) {
        const uint64_t inst = unit * per + team;
        const bool has = member && inst < n;

        // ---- phase 0: stage both operands in LDS, find their significant
        // lengths (to a multiple of 4 digits, enough for the rounding to 8)
        if (has)
            for (uint32_t i = 4 * tl; i < kWinPad; i += 4 * T) *reinterpret_cast<u32x4*>(W + i) = u32x4{0u, 0u, 0u, 0u};
        if (has && tl == 0) {
            META[0] = 0;
            META[1] = 0;
        }
        team_sync<WAVE>();
        if (has) {
            const uint32_t* ga = as + inst * m;
            const uint32_t* gb = bs + inst * m;
            uint32_t pa = 0, pb = 0;
            for (uint32_t i = 4 * tl; i < M8 + 16; i += 4 * T) {
                u32x4 va, vb;
                if (VEC && i + 4 <= m) {
                    va = ld4<true>(ga + i);
                    vb = ld4<true>(gb + i);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        va[j] = (i + j < m) ? ga[i + j] : 0u;
                        vb[j] = (i + j < m) ? gb[i + j] : 0u;
                    }
                }
                if (i < M8 + 12) *reinterpret_cast<u32x4*>(A + i) = va;
                *reinterpret_cast<u32x4*>(W + kWinPad + i) = vb;
                pa = (va[0] | va[1] | va[2] | va[3]) ? i + 4 : pa;
                pb = (vb[0] | vb[1] | vb[2] | vb[3]) ? i + 4 : pb;
            }
            if (pa) atomicMax(&META[0], pa);
            if (pb) atomicMax(&META[1], pb);
        }
        team_sync<WAVE>();

        // ---- phase 1: per-lane block schedule and the convolution
        LaneSched sc = {0, 0, 0, 0, 0, 0};
        uint32_t K = 0;
        if (has) {
            const uint32_t na = META[0], nb = META[1];
            const uint32_t Nf = ((na > nb ? na : nb) + 7) & ~7u;
            if (2 * Nf <= M8) {           // full product of Nf x Nf digits
                sc = sched_rect_r<R>(tl, Nf / 4, Nf / 4);
                K = 2 * Nf;
            } else {                      // low M8 columns of M8 x M8
                sc = sched_low_r<R>(tl, geo.G);
                K = M8;
            }
        }
        const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
        uint64_t acc[4];
        uint32_t cnt[4];
        conv_walk<R>(A, W + 16 - 2 * R, sc, S, snap, acc, cnt);
        team_sync<WAVE>();   // every lane is done with the operands
        if (has) {
            put_lane_groups(E, O, sc, snap, acc, cnt);
            zero_groups(E, O, K, M8, tl, T);
        }
        team_sync<WAVE>();

        // ---- phase 2: recombine column sums into digits, one wave per run of
        // consecutive instances, 256 digits per tile: digit k before carries
        // is E[k] + O[k] < 2^33; its low word and its carry bit (added one
        // digit up) go through one carry look-ahead.
        const uint32_t lane = ln;
        uint32_t t_count = (WAVE || wv == 0) ? per : 0;
        const uint64_t inst0 = unit * per;
        if (inst0 >= n) t_count = 0;
        else if (inst0 + t_count > n) t_count = (uint32_t)(n - inst0);
        const uint32_t ndig = t_count * m;
        uint32_t* out = rs + inst0 * m;
        const bool vec_out = VEC && ((((uintptr_t)out) & 15u) == 0);
        uint32_t carry = 0;
        for (uint32_t q0 = 0; q0 < ndig; q0 += 256) {
            const uint32_t qb = q0 + 4 * lane;
            u32x4 X = {0u, 0u, 0u, 0u}, Y = {0u, 0u, 0u, 0u};
            uint32_t fl = 0;
            if (qb < ndig) rec_digits<VEC>(tlds + 4, geo.words, CW, m, geo.magic, qb, ndig, X, Y, fl);
            u32x4 sv;
            bool g, p;
            rec_lane(X, Y, fl, sv, g, p);
            uint32_t cout;
            const uint32_t e = wave_carry(g, p, carry, cout);
            carry = cout;
            const u32x4 r = rec_post(sv, fl, e);
            if (vec_out && qb + 4 <= ndig) {
                st4<true>(out + qb, r);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (qb + j < ndig) out[qb + j] = r[j];
            }
        }
        team_sync<WAVE>();   // the next group reuses this LDS
    }
}

// Short integers (m <= kMulSmallDigits): one lane per instance, both operands
// in registers, product scanning from the lowest column up -- column k's
// products a_j b_(k-j) go into a 96-bit accumulator (bi::mac: 64-bit sum and
// overflow count), its low word is digit k and the rest carries into column
// k+1.  M = m rounded up to a multiple of 4 (the digits above m read as
// zero), every loop unrolled so the operands stay in registers: M(M+1)/2
// multiply-adds a lane and no team schedule -- at these sizes the team
// kernel's per-instance steps cost more than the multiply-adds, and millions
// of instances keep every lane busy.  Where every instance of a wave has its
// operands' top halves zero (both below B^(M/2)), the wave takes the whole
// product of the low halves instead: (M/2)^2 multiply-adds, the same digits
// (the team kernel's full-product schedule makes the same choice).
//
// A workgroup takes R = small_rows(M) consecutive instances, whose digits are
// one contiguous run of rows*m words in each array.  Each operand moves
// through one LDS tile (rows of M+1 words: an odd stride, so the 64 lanes of
// a wave reading word k of their own rows hit distinct banks): coalesced
// 16-byte (VEC: m == M, 16-byte aligned arrays) or 4-byte loads fill it, each
// lane copies its row to registers; the product digits go back through the
// same tile and leave in coalesced stores.  Reading rows straight from HBM,
// one lane per row, would touch a separate line per lane in every load and
// pull each line through the caches several times.
template <int M>
constexpr uint32_t small_rows() {
    return M <= 32 ? 256u : 128u;   // the tile stays under 64 KiB
}

template <int M, bool VEC>
__device__ __forceinline__ void small_fill(uint32_t* __restrict__ t, const uint32_t* __restrict__ g, uint32_t m,
                                           uint32_t words) {
    constexpr uint32_t S = M + 1, R = small_rows<M>();
    const uint32_t tid = threadIdx.x;
    if (VEC) {
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)M / 4; j++) {
            const uint32_t w = 4 * tid + 4 * R * j;
            if (w < words) {
                const u32x4 x = ld4<true>(g + w);
                uint32_t* d = t + (w / M) * S + (w % M);
                d[0] = x[0];
                d[1] = x[1];
                d[2] = x[2];
                d[3] = x[3];
            }
        }
    } else {
        // word tid + R j: row and column stepped by (R / m, R % m)
        uint32_t r = tid / m, c = tid - (tid / m) * m;
        const uint32_t dq = R / m, dr = R - dq * m;
#pragma unroll 2
        for (uint32_t j = 0; j < (uint32_t)M; j++) {
            const uint32_t w = tid + R * j;
            if (w < words) t[r * S + c] = g[w];
            r += dq;
            c += dr;
            if (c >= m) {
                c -= m;
                r++;
            }
        }
    }
}

template <int M, bool VEC>
__global__ void __launch_bounds__(small_rows<M>()) k_mul_small(const uint32_t* __restrict__ as,
                                                               const uint32_t* __restrict__ bs,
                                                               uint32_t* __restrict__ rs, uint32_t m, uint64_t n) {
    constexpr uint32_t S = M + 1, R = small_rows<M>();
    constexpr int H = M / 2;
    __shared__ uint32_t t[R * S];
    const uint32_t tid = threadIdx.x;
    const uint64_t first = (uint64_t)blockIdx.x * R;
    const uint32_t rows = (uint32_t)min<uint64_t>(R, n - first);
    const uint32_t words = rows * m;
    const uint64_t base = first * m;
    uint32_t a[M], b[M];
    // lanes past the batch's last row read tile words nobody wrote and store nothing
    small_fill<M, VEC>(t, as + base, m, words);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < M; k++) a[k] = (VEC || (uint32_t)k < m) ? t[tid * S + k] : 0u;
    __syncthreads();
    small_fill<M, VEC>(t, bs + base, m, words);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < M; k++) b[k] = (VEC || (uint32_t)k < m) ? t[tid * S + k] : 0u;
    __syncthreads();
    uint32_t top = 0;
#pragma unroll
    for (int k = H; k < M; k++) top |= a[k] | b[k];
    uint64_t lo = 0;
    uint32_t hi = 0;
    if (ballot(top != 0) == 0) {
        // whole product of the low halves: columns 0 .. M-1
#pragma unroll
        for (int k = 0; k < M; k++) {
#pragma unroll
            for (int j = (k - H + 1 > 0 ? k - H + 1 : 0); j <= (k < H - 1 ? k : H - 1); j++)
                mac(lo, hi, a[j], b[k - j]);
            t[tid * S + k] = (uint32_t)lo;
            lo = (lo >> 32) | ((uint64_t)hi << 32);
            hi = 0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < M; k++) {
#pragma unroll
            for (int j = 0; j <= k; j++) mac(lo, hi, a[j], b[k - j]);
            t[tid * S + k] = (uint32_t)lo;
            lo = (lo >> 32) | ((uint64_t)hi << 32);
            hi = 0;
        }
    }
    __syncthreads();
    uint32_t* __restrict__ g = rs + base;
    if (VEC) {
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)M / 4; j++) {
            const uint32_t w = 4 * tid + 4 * R * j;
            if (w < words) {
                const uint32_t* d = t + (w / M) * S + (w % M);
                st4<true>(g + w, u32x4{d[0], d[1], d[2], d[3]});
            }
        }
    } else {
        uint32_t r = tid / m, c = tid - (tid / m) * m;
        const uint32_t dq = R / m, dr = R - dq * m;
#pragma unroll 2
        for (uint32_t j = 0; j < (uint32_t)M; j++) {
            const uint32_t w = tid + R * j;
            if (w < words) g[w] = t[r * S + c];
            r += dq;
            c += dr;
            if (c >= m) {
                c -= m;
                r++;
            }
        }
    }
}

hipError_t mul_uniform(const MulCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                       uint64_t n, hipStream_t st, uint32_t* snaps, uint64_t snap_words, uint32_t num_cus) {
    if (n == 0) return hipSuccess;
    if (m <= kMulSmallDigits && n >= (m <= 32 ? kMulSmallMinInstances : kMulSmallMinInstancesWide)) {
        const bool vec = (m % 4 == 0) && ((((uintptr_t)as) | ((uintptr_t)bs) | ((uintptr_t)rs)) & 15u) == 0;
        // at most 2^22 workgroups a launch (grid x dimension times block size below 2^32)
        const uint64_t per = (uint64_t)128 << 22;
        for (uint64_t i0 = 0; i0 < n; i0 += per) {
            const uint64_t cnt = std::min<uint64_t>(per, n - i0);
            const uint32_t* a0 = as + i0 * m;
            const uint32_t* b0 = bs + i0 * m;
            uint32_t* r0 = rs + i0 * m;
#define BI_SMALL(M)                                                                                               \
    do {                                                                                                          \
        constexpr uint32_t R = small_rows<M>();                                                                   \
        const dim3 grid((uint32_t)((cnt + R - 1) / R)), block(R);                                                 \
        if (vec) hipLaunchKernelGGL((k_mul_small<M, true>), grid, block, 0, st, a0, b0, r0, m, cnt);              \
        else hipLaunchKernelGGL((k_mul_small<M, false>), grid, block, 0, st, a0, b0, r0, m, cnt);                 \
    } while (0)
        switch ((m + 3) / 4) {
            case 1: BI_SMALL(4); break;
            case 2: BI_SMALL(8); break;
            case 3: BI_SMALL(12); break;
            case 4: BI_SMALL(16); break;
            case 5: BI_SMALL(20); break;
            case 6: BI_SMALL(24); break;
            case 7: BI_SMALL(28); break;
            case 8: BI_SMALL(32); break;
            case 9: BI_SMALL(36); break;
            case 10: BI_SMALL(40); break;
            case 11: BI_SMALL(44); break;
            case 12: BI_SMALL(48); break;
            case 13: BI_SMALL(52); break;
            case 14: BI_SMALL(56); break;
            case 15: BI_SMALL(60); break;
            default: BI_SMALL(64); break;
        }
#undef BI_SMALL
        }
        return hipGetLastError();
    }
    if (cfg.limb_bits == 8 && mul_mfma_eligible(m)) return mul_mfma_uniform(as, bs, rs, m, n, st, num_cus);
    if (cfg.limb_bits == 27 && mul27_eligible(m)) return mul27_uniform(as, bs, rs, m, n, st, snaps, snap_words, num_cus);
    uint32_t wpb = cfg.waves_per_block ? cfg.waves_per_block : 4;
    MulGeom g = mul_geom(m, wpb);
    while ((size_t)g.teams_per_block * g.words * 4 > 160 * 1024 && wpb > 1) g = mul_geom(m, wpb /= 2);
    const size_t lds = (size_t)g.teams_per_block * g.words * 4;
    const bool wave = g.tpw != 0;
    const uint64_t per = wave ? g.tpw : 1;
    const uint64_t nunits = (n + per - 1) / per;
    const uint64_t units_per_block = wave ? g.block / 64 : 1;
    const bool vec = (m % 4 == 0) && ((((uintptr_t)as) | ((uintptr_t)bs) | ((uintptr_t)rs)) & 15u) == 0;
    static const bool attr_done = [] {   // once per process, thread-safe
        set_max_lds(reinterpret_cast<const void*>(&k_mul<true, 4, true>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<false, 4, true>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<true, 8, true>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<false, 8, true>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<true, 4, false>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<false, 4, false>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<true, 8, false>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul<false, 8, false>), 160 * 1024);
        return true;
    }();
    (void)attr_done;
    typedef void (*Kernel)(const uint32_t*, const uint32_t*, uint32_t*, uint64_t, MulGeom, uint32_t*, uint64_t);
    Kernel fn;
    if (cfg.rows == 4) fn = wave ? (vec ? k_mul<true, 4, true> : k_mul<false, 4, true>)
                                 : (vec ? k_mul<true, 4, false> : k_mul<false, 4, false>);
    else fn = wave ? (vec ? k_mul<true, 8, true> : k_mul<false, 8, true>)
                   : (vec ? k_mul<true, 8, false> : k_mul<false, 8, false>);
    // persistent grid: the workgroups resident on the device at once
    const uint32_t per_cu = resident_per_cu(reinterpret_cast<const void*>(fn), g.block, lds);
    uint64_t grid = std::min<uint64_t>((nunits + units_per_block - 1) / units_per_block, (uint64_t)num_cus * per_cu);
    if (grid * g.block * 12 > snap_words) grid = snap_words / (12ull * g.block);
    if (grid == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3((uint32_t)grid), dim3(g.block), lds, st, as, bs, rs, n, g, snaps, nunits);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Chained products kept in LDS: rs = (as*bs)^p mod B^m, computed as the
// reference's bench chain r1 = a*b, r_i = r1 * r_{i-1} (cuda/ker-mul.cu.h:
// 223-250 convMul1Bench and its v2..v5 twins).  Teams of <= 64 lanes only
// (m <= 512): a wave keeps the result digits of its teams in registers (at
// most 512 digits = 2 tiles), restages r1 and r_i as the next operands and
// runs the next product without touching HBM.
template <bool VEC, int R>
__global__ void __launch_bounds__(1024) k_mul_chain(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                    uint32_t* __restrict__ rs, uint64_t n, MulGeom geo,
                                                    uint32_t* __restrict__ snaps, uint64_t nunits, uint32_t p) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t m = geo.m, M8 = geo.M8, T = geo.T, CW = geo.CW;
    const uint32_t wv = tid / 64, ln = tid % 64;
    const uint32_t team = ln / T, tl = ln % T, per = geo.tpw;
    const bool member = ln < geo.tpw * T;
    uint32_t* tlds = lds + (size_t)wv * geo.tpw * geo.words;
    const uint64_t u0 = (uint64_t)blockIdx.x * (blockDim.x / 64) + wv;
    const uint64_t ustride = (uint64_t)gridDim.x * (blockDim.x / 64);
    uint32_t* META = tlds + (size_t)team * geo.words;
    uint32_t* U = META + 4;
    uint32_t* A = U;
    uint32_t* W = U + M8 + 12;
    uint32_t* E = U;
    uint32_t* O = U + CW;
    uint32_t* snap = snaps + ((size_t)blockIdx.x * blockDim.x + tid) * 12;

    for (uint64_t unit = u0; unit < nunits; unit += ustride) {
        const uint64_t inst = unit * per + team;
        const bool has = member && inst < n;
        uint32_t t_count = per;
        const uint64_t inst0 = unit * per;
        if (inst0 + t_count > n) t_count = (uint32_t)(n - inst0);
        const uint32_t ndig = t_count * m;

        // stage a and b, significant lengths (first product only)
        if (has) {
            for (uint32_t i = 4 * tl; i < kWinPad; i += 4 * T) *reinterpret_cast<u32x4*>(W + i) = u32x4{0u, 0u, 0u, 0u};
            if (tl == 0) {
                META[0] = 0;
                META[1] = 0;
            }
        }
        team_sync<true>();
        if (has) {
            const uint32_t* ga = as + inst * m;
            const uint32_t* gb = bs + inst * m;
            uint32_t pa = 0, pb = 0;
            for (uint32_t i = 4 * tl; i < M8 + 16; i += 4 * T) {
                u32x4 va, vb;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    va[j] = (i + j < m) ? ga[i + j] : 0u;
                    vb[j] = (i + j < m) ? gb[i + j] : 0u;
                }
                if (i < M8 + 12) *reinterpret_cast<u32x4*>(A + i) = va;
                *reinterpret_cast<u32x4*>(W + kWinPad + i) = vb;
                pa = (va[0] | va[1] | va[2] | va[3]) ? i + 4 : pa;
                pb = (vb[0] | vb[1] | vb[2] | vb[3]) ? i + 4 : pb;
            }
            if (pa) atomicMax(&META[0], pa);
            if (pb) atomicMax(&META[1], pb);
        }
        team_sync<true>();

        u32x4 r1v[2], rv[2];   // this lane's digits of tiles 0, 1 of r1 and of r_i
        for (uint32_t it = 0; it < p; it++) {
            if (it > 0) {
                // restage A = r1, W = r_i (digits of instance tm at k)
                if (has) {
                    for (uint32_t i = m + tl; i < M8 + 12; i += T) A[i] = 0u;
                    for (uint32_t i = tl; i < kWinPad; i += T) W[i] = 0u;
                    for (uint32_t i = m + tl; i < M8 + 16; i += T) W[kWinPad + i] = 0u;
                }
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const uint32_t qb = 256 * q + 4 * ln;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const uint32_t d = qb + j;
                        if (d < ndig) {
                            const uint32_t tm = (uint32_t)(((uint64_t)d * geo.magic) >> 32);
                            const uint32_t k = d - tm * m;
                            uint32_t* tb = tlds + (size_t)tm * geo.words + 4;
                            tb[k] = r1v[q][j];
                            tb[M8 + 12 + kWinPad + k] = rv[q][j];
                        }
                    }
                }
                team_sync<true>();
            }
            LaneSched sc = {0, 0, 0, 0, 0, 0};
            uint32_t K = 0;
            if (has) {
                const uint32_t na = it ? M8 : META[0], nb = it ? M8 : META[1];
                const uint32_t Nf = ((na > nb ? na : nb) + 7) & ~7u;
                if (2 * Nf <= M8) {
                    sc = sched_rect_r<R>(tl, Nf / 4, Nf / 4);
                    K = 2 * Nf;
                } else {
                    sc = sched_low_r<R>(tl, geo.G);
                    K = M8;
                }
            }
            const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
            uint64_t acc[4];
            uint32_t cnt[4];
            conv_walk<R>(A, W + 16 - 2 * R, sc, S, snap, acc, cnt);
            team_sync<true>();
            if (has) {
                put_lane_groups(E, O, sc, snap, acc, cnt);
                zero_groups(E, O, K, M8, tl, T);
            }
            team_sync<true>();
            // recombine into registers
            uint32_t carry = 0;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const uint32_t qb = 256 * q + 4 * ln;
                if (256 * q >= ndig) break;
                u32x4 X = {0u, 0u, 0u, 0u}, Y = {0u, 0u, 0u, 0u};
                uint32_t fl = 0;
                if (qb < ndig) rec_digits<false>(tlds + 4, geo.words, CW, m, geo.magic, qb, ndig, X, Y, fl);
                u32x4 sv;
                bool g, pp;
                rec_lane(X, Y, fl, sv, g, pp);
                uint32_t cout;
                const uint32_t e = wave_carry(g, pp, carry, cout);
                carry = cout;
                rv[q] = rec_post(sv, fl, e);
                if (it == 0) r1v[q] = rv[q];
            }
            team_sync<true>();   // column reads done before the next restage
        }
        // store r_p
        uint32_t* out = rs + inst0 * m;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t qb = 256 * q + 4 * ln;
            if (VEC && qb + 4 <= ndig) {
                st4<true>(out + qb, rv[q]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (qb + j < ndig) out[qb + j] = rv[q][j];
            }
        }
        team_sync<true>();
    }
}

hipError_t mul_chain_uniform(const MulCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                             uint64_t n, uint32_t p, hipStream_t st, uint32_t* snaps, uint64_t snap_words,
                             uint32_t num_cus) {
    if (n == 0) return hipSuccess;
    if (m > kMulChainMaxDigits || p == 0) return hipErrorInvalidValue;
    uint32_t wpb = cfg.waves_per_block ? cfg.waves_per_block : 4;
    MulGeom g = mul_geom(m, wpb);
    while ((size_t)g.teams_per_block * g.words * 4 > 160 * 1024 && wpb > 1) g = mul_geom(m, wpb /= 2);
    if (!g.tpw) return hipErrorInvalidValue;
    const size_t lds = (size_t)g.teams_per_block * g.words * 4;
    const uint64_t nunits = (n + g.tpw - 1) / g.tpw;
    const uint64_t units_per_block = g.block / 64;
    const bool vec = (m % 4 == 0) && ((((uintptr_t)rs) & 15u) == 0);
    static const bool attr_done = [] {   // once per process, thread-safe
        set_max_lds(reinterpret_cast<const void*>(&k_mul_chain<true, 4>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul_chain<false, 4>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul_chain<true, 8>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_mul_chain<false, 8>), 160 * 1024);
        return true;
    }();
    (void)attr_done;
    typedef void (*Kernel)(const uint32_t*, const uint32_t*, uint32_t*, uint64_t, MulGeom, uint32_t*, uint64_t,
                           uint32_t);
    const Kernel fn = cfg.rows == 4 ? (vec ? k_mul_chain<true, 4> : k_mul_chain<false, 4>)
                                    : (vec ? k_mul_chain<true, 8> : k_mul_chain<false, 8>);
    const uint32_t per_cu = resident_per_cu(reinterpret_cast<const void*>(fn), g.block, lds);
    uint64_t grid = std::min<uint64_t>((nunits + units_per_block - 1) / units_per_block, (uint64_t)num_cus * per_cu);
    if (grid * g.block * 12 > snap_words) grid = snap_words / (12ull * g.block);
    if (grid == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3((uint32_t)grid), dim3(g.block), lds, st, as, bs, rs, n, g, snaps, nunits, p);
    return hipGetLastError();
}

uint64_t mul_snap_words(uint32_t num_cus) { return (uint64_t)num_cus * 2048 * 12; }

}  // namespace bi
