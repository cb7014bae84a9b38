// bi_mul27.hip -- batched multiply through 27-bit limbs: rs = low m digits of
// as*bs per instance, for m <= kMul27MaxDigits.  Same results as the 32-bit
// convolution of bi_mul.hip (reference semantics: cuda/ker-mul.cu.h:147-666,
// the low m digits; cuda/helpers.h:161-187 truncates the GMP product to m
// words); a different way to reach them.
//
// Why 27 bits.  The 32-bit convolution spends two VALU instructions per digit
// product: v_mad_u64_u32 into a 64-bit column sum and one add that counts the
// column's overflows.  With limbs of 27 bits a product is below 2^54, and a
// column of at most 1024 products stays below 2^64: the column sums need no
// overflow count, and a limb product costs the one v_mad_u64_u32.  There are
// (32/27)^2 = 1.40 times as many limb products as digit products, so the
// multiply-add stream shrinks to 1.40/2 = 0.70 of the 32-bit one.  1024
// limbs = 864 digits bounds m.
//
// Per instance (team of L8/8 lanes, L8 = limbs of m digits rounded up to 8):
//   stage   limb t = bits [27t, 27t+27) of each operand, read straight from
//           HBM (two digits per limb, v_alignbit), into the team's LDS with the
//           operand layout of bi_mulcore.h (limbs in place of digits); the
//           significant limbs of both operands decide the schedule;
//   walk    the R-row block schedules of bi_mulcore.h over limbs (truncated:
//           columns < L8; full, when both operands fit in Lf <= L8/2 limbs:
//           columns < 2Lf only), one v_mad_u64_u32 per limb product;
//   fold    a lane folds each of its 4-column groups into six 27-bit limbs
//           (sum of col_{k+j} * 2^(27j), k = first column of the group) and
//           stores them 8 words wide, even groups in E, odd groups in O, so
//           limb t before carries is E[t] + O[t] < 2^28;
//   carry   one wave per instance: a carry look-ahead in base 2^27 (carries
//           0 or 1) normalises the limbs in place, then each output digit is
//           bits [32d, 32d+32) of the limbs (three limbs, one 64-bit shift).
//
// Teams of 76 lanes (m = 512) do not divide a wave, so teams are laid out
// across the waves of a workgroup (e.g. five teams in six waves) and the
// phases are separated by workgroup barriers.  That is the price of the
// scheme: the staging loads and the carry pass of a workgroup no longer
// overlap other waves' walks as they do in the wave-per-instance 32-bit
// kernel, so the limb convolution wins only where its teams fill waves
// (m = 864: 128 lanes); the autotuner chooses per shape.
#include "bi_kernels.h"
#include "bi_mulcore.h"

#include <algorithm>

namespace bi {

constexpr uint32_t kLimbBits = 27;
constexpr uint32_t kLimbMask = (1u << kLimbBits) - 1;

__host__ __device__ inline uint32_t limbs_of(uint32_t digits) { return (32 * digits + kLimbBits - 1) / kLimbBits; }

struct Mul27Geom {
    uint32_t m;      // digits per instance
    uint32_t L8;     // limbs of m digits, rounded up to 8
    uint32_t T;      // lanes per team (L8 / 8)
    uint32_t tpb;    // teams (instances) per workgroup
    uint32_t block;  // threads per workgroup (whole waves)
    uint32_t words;  // LDS words per team: META[4] | operands or column arrays
    uint32_t CW;     // column array stride (words)
};

// Limb t of the m-digit integer x: bits [27t, 27t+27), 0 above the top.
__device__ __forceinline__ uint32_t limb_at(const uint32_t* __restrict__ x, uint32_t m, uint32_t t) {
    const uint32_t bit = kLimbBits * t, j = bit >> 5;
    if (j >= m) return 0u;
    const uint32_t lo = x[j];
    const uint32_t hi = j + 1 < m ? x[j + 1] : 0u;
    return __builtin_amdgcn_alignbit(hi, lo, bit & 31u) & kLimbMask;
}

// acc[c] += a * w_c: four v_mad_u64_u32, no overflow possible (see above).
__device__ __forceinline__ void mad4(uint64_t* acc, uint32_t a, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
    acc[0] += (uint64_t)a * w0;
    acc[1] += (uint64_t)a * w1;
    acc[2] += (uint64_t)a * w2;
    acc[3] += (uint64_t)a * w3;
}

// conv_walk of bi_mulcore.h over limbs: the same pointers and zero-window
// saturation, 64-bit sums only; the snapshot is the first group's four sums
// (8 words of the lane's global scratch).
template <int R>
__device__ __forceinline__ void walk27(const uint32_t* A, const uint32_t* Wr, const LaneSched sc, uint32_t S,
                                       uint32_t* snap, uint64_t* acc) {
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
    for (int c = 0; c < 4; c++) acc[c] = 0;
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
        for (int r = 0; r < R; r++) mad4(acc, av[r], wv[R - r], wv[R + 1 - r], wv[R + 2 - r], wv[R + 3 - r]);
        pa += R;
        pb = __builtin_elementwise_sub_sat(pb, (uint32_t)R);
        if (s + 1 == sw) {
            *reinterpret_cast<u32x4*>(snap) =
                u32x4{(uint32_t)acc[0], (uint32_t)(acc[0] >> 32), (uint32_t)acc[1], (uint32_t)(acc[1] >> 32)};
            *reinterpret_cast<u32x4*>(snap + 4) =
                u32x4{(uint32_t)acc[2], (uint32_t)(acc[2] >> 32), (uint32_t)acc[3], (uint32_t)(acc[3] >> 32)};
            pa = pa2;
            pb = pb2;
        }
    }
}

// The four column sums of the group at limb column k (k % 4 == 0) folded into
// six 27-bit limbs, stored with 2 zero words at E[k..k+8) (k % 8 == 0) or
// O[k..k+8).  The running value stays below 2^65: its 65th bit is kept as o.
__device__ __forceinline__ void put_group27(uint32_t* E, uint32_t* O, uint32_t k, const uint64_t* acc) {
    uint32_t l[6];
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint64_t x = acc[j] + c;
        const uint64_t o = x < c;
        l[j] = (uint32_t)x & kLimbMask;
        c = (x >> kLimbBits) | (o << (64 - kLimbBits));
    }
    l[4] = (uint32_t)c & kLimbMask;
    l[5] = (uint32_t)(c >> kLimbBits);
    uint32_t* dst = ((k & 7u) ? O : E) + k;
    *reinterpret_cast<u32x4*>(dst) = u32x4{l[0], l[1], l[2], l[3]};
    *reinterpret_cast<u32x4*>(dst + 4) = u32x4{l[4], l[5], 0u, 0u};
}

__device__ __forceinline__ void put_lane_groups27(uint32_t* E, uint32_t* O, const LaneSched sc, const uint32_t* snap,
                                                  uint64_t* acc) {
    if (sc.stepsA + sc.stepsB == 0) return;
    if (sc.stepsB) {
        const u32x4 s0 = *reinterpret_cast<const u32x4*>(snap);
        const u32x4 s1 = *reinterpret_cast<const u32x4*>(snap + 4);
        uint64_t ra[4] = {((uint64_t)s0[1] << 32) | s0[0], ((uint64_t)s0[3] << 32) | s0[2],
                          ((uint64_t)s1[1] << 32) | s1[0], ((uint64_t)s1[3] << 32) | s1[2]};
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] -= ra[j];
        put_group27(E, O, sc.kA, ra);
        put_group27(E, O, sc.kB, acc);
    } else {
        put_group27(E, O, sc.kA, acc);
    }
}

// Output digit d (bits [32d, 32d+32)) of the normalised limbs N[0..L8).
__device__ __forceinline__ uint32_t digit_of(const uint32_t* N, uint32_t L8, uint32_t d) {
    const uint32_t bit = 32 * d, t = bit / kLimbBits, b = bit - kLimbBits * t;
    const uint64_t n0 = N[t];
    const uint64_t n1 = t + 1 < L8 ? N[t + 1] : 0u;
    const uint64_t n2 = t + 2 < L8 ? N[t + 2] : 0u;
    const uint64_t x = n0 | (n1 << kLimbBits) | (n2 << (2 * kLimbBits));
    return (uint32_t)(x >> b);
}

template <bool VEC, int R>
__global__ void __launch_bounds__(1024) k_mul27(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                uint32_t* __restrict__ rs, uint64_t n, Mul27Geom geo,
                                                uint32_t* __restrict__ snaps, uint64_t nunits) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t m = geo.m, L8 = geo.L8, T = geo.T, CW = geo.CW, tpb = geo.tpb;
    const uint32_t team = tid / T, tl = tid % T;
    const bool member = team < tpb;
    uint32_t* META = lds + (size_t)(member ? team : 0) * geo.words;   // [0] limbs of a, [1] limbs of b
    uint32_t* U = META + 4;
    uint32_t* A = U;
    uint32_t* W = U + L8 + 12;
    uint32_t* E = U;
    uint32_t* O = U + CW;
    uint32_t* snap = snaps + ((size_t)blockIdx.x * blockDim.x + tid) * 8;
    const uint32_t wv = tid / 64, lane = tid % 64, nw = blockDim.x / 64;

    if (member && tl == 0) {
        META[0] = 0;
        META[1] = 0;
    }
    __syncthreads();
    for (uint64_t unit = blockIdx.x; unit < nunits; unit += gridDim.x) {
        const uint64_t inst = unit * tpb + team;
        const bool has = member && inst < n;

        // ---- stage both operands as limbs, find their significant limbs
        if (has) {
            const uint32_t* ga = as + inst * m;
            const uint32_t* gb = bs + inst * m;
            uint32_t la = 0, lb = 0;
            for (uint32_t t = tl; t < L8 + 16; t += T) {
                const uint32_t va = limb_at(ga, m, t), vb = limb_at(gb, m, t);
                if (t < L8 + 12) A[t] = va;
                W[kWinPad + t] = vb;
                la = va ? t + 1 : la;
                lb = vb ? t + 1 : lb;
            }
            for (uint32_t i = tl; i < kWinPad; i += T) W[i] = 0u;
            if (la) atomicMax(&META[0], la);
            if (lb) atomicMax(&META[1], lb);
        }
        __syncthreads();

        // ---- lane schedule and the walk
        LaneSched sc = {0, 0, 0, 0, 0, 0};
        uint32_t K = 0;
        if (has) {
            const uint32_t na = META[0], nb = META[1];
            const uint32_t Lf = ((na > nb ? na : nb) + 7) & ~7u;
            if (2 * Lf <= L8) {           // full product of Lf x Lf limbs
                sc = sched_rect_r<R>(tl, Lf / 4, Lf / 4);
                K = 2 * Lf;
            } else {                      // low L8 columns of L8 x L8
                sc = sched_low_r<R>(tl, L8 / 4);
                K = L8;
            }
        }
        const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
        uint64_t acc[4];
        walk27<R>(A, W + 16 - 2 * R, sc, S, snap, acc);
        __syncthreads();   // every lane is done with the operands
        if (has) {
            put_lane_groups27(E, O, sc, snap, acc);
            zero_groups(E, O, K, L8, tl, T);
            if (tl == 0) {   // read before the barrier above; the next unit's staging counts afresh
                META[0] = 0;
                META[1] = 0;
            }
        }
        __syncthreads();

        // ---- carries: one wave per instance; limb t = E[t] + O[t] < 2^28
        for (uint32_t q = wv; q < tpb; q += nw) {
            const uint64_t qi = unit * tpb + q;
            if (qi >= n) break;
            uint32_t* Eq = lds + (size_t)q * geo.words + 4;
            const uint32_t* Oq = Eq + CW;
            uint32_t carry = 0;
            for (uint32_t t0 = 0; t0 < L8; t0 += 256) {
                const uint32_t t = t0 + 4 * lane;   // L8 % 8 == 0: a lane's 4 limbs are all in or all out
                u32x4 v = {0u, 0u, 0u, 0u};
                if (t < L8) v = *reinterpret_cast<const u32x4*>(Eq + t) + *reinterpret_cast<const u32x4*>(Oq + t);
                uint32_t c = 0, all = kLimbMask;
                u32x4 s;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t x = v[j] + c;
                    s[j] = x & kLimbMask;
                    c = x >> kLimbBits;
                    all &= s[j];
                }
                uint32_t cout;
                uint32_t e = wave_carry(c != 0, all == kLimbMask, carry, cout);
                carry = cout;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t x = s[j] + e;
                    s[j] = x & kLimbMask;
                    e = x >> kLimbBits;
                }
                if (t < L8) *reinterpret_cast<u32x4*>(Eq + t) = s;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            uint32_t* out = rs + qi * m;
            const bool vec_out = VEC && ((((uintptr_t)out) & 15u) == 0);
            for (uint32_t d0 = 0; d0 < m; d0 += 256) {
                const uint32_t d = d0 + 4 * lane;
                u32x4 r;
#pragma unroll
                for (int j = 0; j < 4; j++) r[j] = d + j < m ? digit_of(Eq, L8, d + j) : 0u;
                if (vec_out && d + 4 <= m) {
                    st4<true>(out + d, r);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (d + j < m) out[d + j] = r[j];
                }
            }
        }
        __syncthreads();   // the next unit reuses this LDS
    }
}

// Workgroup shape: the number of waves whose lanes the teams fill best (ties:
// fewer waves), within 64 KB of LDS so that several workgroups share a CU.
static Mul27Geom mul27_geom(uint32_t m) {
    Mul27Geom g;
    g.m = m;
    g.L8 = (limbs_of(m) + 7) & ~7u;
    g.T = g.L8 / 8;
    g.CW = team_col_stride(g.L8);
    g.words = team_words(g.L8, g.L8);
    g.tpb = 0;
    g.block = 0;
    double best = -1;
    for (uint32_t w = 1; w <= 16; w++) {
        const uint32_t tpb = 64 * w / g.T;
        if (tpb == 0 || (size_t)tpb * g.words * 4 > 64 * 1024) continue;
        const double util = (double)tpb * g.T / (64.0 * w);
        if (util > best + 1e-9) {
            best = util;
            g.tpb = tpb;
            g.block = 64 * w;
        }
    }
    return g;
}

bool mul27_eligible(uint32_t m) { return m >= 1 && m <= kMul27MaxDigits; }

hipError_t mul27_uniform(const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m, uint64_t n, hipStream_t st,
                         uint32_t* snaps, uint64_t snap_words, uint32_t num_cus) {
    if (n == 0) return hipSuccess;
    if (!mul27_eligible(m)) return hipErrorInvalidValue;
    const Mul27Geom g = mul27_geom(m);
    if (g.tpb == 0) return hipErrorInvalidValue;
    const size_t lds = (size_t)g.tpb * g.words * 4;
    const uint64_t nunits = (n + g.tpb - 1) / g.tpb;
    const bool vec = (m % 4 == 0) && ((((uintptr_t)rs) & 15u) == 0);
    typedef void (*Kernel)(const uint32_t*, const uint32_t*, uint32_t*, uint64_t, Mul27Geom, uint32_t*, uint64_t);
    const Kernel fn = vec ? k_mul27<true, 8> : k_mul27<false, 8>;
    // persistent grid: the workgroups resident on the device at once
    const uint32_t per_cu = resident_per_cu(reinterpret_cast<const void*>(fn), g.block, lds);
    uint64_t grid = std::min<uint64_t>(nunits, (uint64_t)num_cus * per_cu);
    if (grid * g.block * 8 > snap_words) grid = snap_words / (8ull * g.block);
    if (grid == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fn, dim3((uint32_t)grid), dim3(g.block), lds, st, as, bs, rs, n, g, snaps, nunits);
    return hipGetLastError();
}

}  // namespace bi
