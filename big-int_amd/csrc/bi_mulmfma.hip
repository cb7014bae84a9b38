// bi_mulmfma.hip -- batched multiply on the int8 matrix cores (gfx950
// v_mfma_i32_32x32x32_i8): rs = low m digits of as*bs per instance, the same
// result as the multiply-add kernels of bi_mul.hip (reference semantics:
// cuda/ker-mul.cu.h:147-666, low m digits of the product, checked against
// GMP truncated to m words, cuda/helpers.h:161-187).
//
// Bytes as signed digits.  The matrix cores multiply signed bytes, so each
// operand x (p significant u32 digits) is rewritten in balanced base 256:
// y = x + K with K = 0x80 in each of its 4p low bytes (one carry look-ahead
// add), then d_i = y_i - 128 (y_i XOR 0x80) for i < 4p and d_4p = the carry
// out of y.  Then x = sum d_i 256^i with every d_i in [-128, 127] and at most
// 4p+1 of them nonzero.
//
// The product as matrix products.  Cut both digit strings into 32-byte blocks
// (a'_p, block p = bytes 32p..32p+31).  Column sums c_k = sum_i a'_i b'_(k-i)
// of the byte product, for k = 32u + t (t < 32), are
//     c_(32u+t) = sum_q sum_i H_q[t][i] a'[32(u-q) + i],
//     H_q[t][i] = b'[32q + t - i]          (a 32x32 Toeplitz block of b'),
// so a 32x32 tile of column sums (t rows; u = 32T..32T+31 columns, 1024
// product bytes = 256 digits) is  sum over q of  H_q x [a'_(u-q)]_u,  one
// MFMA per q accumulating into the same tile.  The MFMA's K index k pairs
// with i = 31 - k, so H_q's row is a run of 16 consecutive bytes of b' per
// lane (read from LDS at any byte offset: 5 dwords and v_alignbyte) and the
// block operand is a' stored byte-reversed in LDS (one aligned 16-byte read).
// Columns whose block index falls outside a' read a zero block.
//
// One wave per instance: stage both operands (signed digits) in LDS, run the
// q loop over the tiles of the low m digits (tiles and q ranges cut to the
// operands' significant digits), then turn the int32 column sums into digits:
// lane (u, h) of tile T holds, in accumulator registers 4g..4g+3, the byte
// sums of digit j = 256T + 8u + 2g + h, so V_j = sum_e c_(4j+e) 2^(8e) is
// formed in registers (|c| < 2^14 * 4m, |V_j| < 2^51 for m <= 1024).  V_j =
// lo_j + 2^32 hi_j; digit j before carries is lo_j + hi_(j-1), which leaves a
// carry of -1, 0 or +1 into digit j+1: one carry look-ahead add of the +1
// carries and one subtract of the -1 carries finish the digits (256 per wave
// pass, bi::wave_carry).
#include "../../include/bigint.h"
#include "bi_addcore.h"
#include "bi_ctx.h"
#include "bi_kernels.h"

namespace bi {

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kMfmaWaves = 4;   // waves per workgroup, each on its own instances


__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS words of one wave (NB = 32*NC blocks of 32 bytes):
//   RA (16*NB + 8): a' byte-reversed inside 2*NB + 1 blocks -- block p
//       (p = -NB-1 .. NB-1) at words (NB-1-p)*8, so digit k of a' sits at
//       word 8NB-1-k; every block outside a' is zero (written once per
//       launch), and the block of any column at any step is read without a
//       range check;
//   BP (M8 + 36): 8 zero words, b', 28 zero words (the operand reads run up
//       to two q steps past the last product step).
// After the products RA[8NB - M8 + j] holds lo_j (the words of a') and
// BP[1 + j] holds hi_j (BP[0] = 0 stands for hi_(-1)).
__host__ __device__ inline uint32_t mfma_wave_words(uint32_t NC, uint32_t M8) { return 512 * NC + 8 + M8 + 36; }

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

// Digits [k, k+4) of x, zero at and above m.
template <bool VEC>
__device__ __forceinline__ u32x4 load4(const uint32_t* __restrict__ x, uint32_t k, uint32_t m) {
    u32x4 v = {0u, 0u, 0u, 0u};
    if (VEC && k + 4 <= m) {
        v = ld4<true>(x + k);
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (k + j < m) v[j] = x[k + j];
    }
    return v;
}

// index + 1 of the highest nonzero digit of x (NC chunks of 4 digits per
// lane), on every lane: in the highest chunk with a nonzero digit, the
// highest lane holding one (ballot) reports its own top digit
template <int NC>
__device__ __forceinline__ uint32_t top_digit(const u32x4 (&x)[NC], uint32_t lane) {
#pragma unroll
    for (int c = NC - 1; c >= 0; c--) {
        uint32_t top = 0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (x[c][j]) top = 256 * c + 4 * lane + j + 1;
        const uint64_t any = ballot(top != 0);
        if (any) return __builtin_amdgcn_readlane(top, 63 - (uint32_t)__builtin_clzll(any));
    }
    return 0;
}

// x (NC chunks of 4 digits per lane) -> its balanced signed bytes: y = x +
// K (K = 0x80 in every byte of the NC*256 digits, one carry look-ahead add),
// then y XOR K.  Above the significant digits this gives the carry digit,
// then zeros, exactly as a K cut to the significant digits would; a carry
// out of the top is a multiple of B^m and drops out of the low m digits.
template <int NC>
__device__ __forceinline__ void to_signed(u32x4 (&x)[NC]) {
    constexpr uint32_t K = 0x80808080u;
    uint32_t carry = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        u32x4 s;
        bool g, p;
        lane_pre<false>(x[c], u32x4{K, K, K, K}, 0u, s, g, p);
        uint32_t cout;
        const uint32_t e = wave_carry(g, p, carry, cout);
        carry = cout;
        const u32x4 y = lane_post<false>(s, 0u, e);
#pragma unroll
        for (int j = 0; j < 4; j++) x[c][j] = y[j] ^ K;
    }
}

// d + y (SUB: d - y) for 0/1 increments y, carries (borrows) local to the
// lane's four digits: the same contract as bi::lane_pre with no instance
// starts (g: carry out of the lane, p: a carry in would pass through all
// four), in 32-bit operations.
template <bool SUB>
__device__ __forceinline__ u32x4 lane_step(const u32x4 d, const u32x4 y, bool& g, bool& p) {
    u32x4 s;
    uint32_t c = 0, all = SUB ? 0u : 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t t = y[j] + c;   // 0, 1 or 2
        uint32_t r;
        c = SUB ? __builtin_sub_overflow(d[j], t, &r) : __builtin_add_overflow(d[j], t, &r);
        s[j] = r;
        if (SUB) all |= r;
        else all &= r;
    }
    g = c != 0;
    p = SUB ? all == 0u : all == 0xFFFFFFFFu;
    return s;
}

// Row t, K half h of H_q: bytes b'[32q + t - 31 + 16h + j], j < 16 (b' at
// byte 32 of BP, so from byte 32q + t + 16h + 1): five dwords, then a byte
// shift of (t + 1) & 3 (rd_h / fin_h, split so that the reads of the next
// step are in flight during this step's products).
struct HRaw {
    uint32_t d[5];
};
// hsrc = BP + ((t + 16h + 1) >> 2): the lane's first word at q = 0; step q
// starts 8q words later (32q bytes, a whole number of words)
__device__ __forceinline__ HRaw rd_h(const uint32_t* hsrc, uint32_t q) {
    const uint32_t* src = hsrc + 8 * q;
    return HRaw{{src[0], src[1], src[2], src[3], src[4]}};
}
__device__ __forceinline__ i32x4 fin_h(const HRaw& r, uint32_t sh) {
    return i32x4{(int)__builtin_amdgcn_alignbyte(r.d[1], r.d[0], sh), (int)__builtin_amdgcn_alignbyte(r.d[2], r.d[1], sh),
                 (int)__builtin_amdgcn_alignbyte(r.d[3], r.d[2], sh), (int)__builtin_amdgcn_alignbyte(r.d[4], r.d[3], sh)};
}

// NC: 256-digit chunks (and 32x32 column-sum tiles) per instance, m <= 256*NC.
template <int NC, bool VEC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NC <= 1 ? 4 : NC == 2 ? 3 : 2))) k_mul_mfma(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                  uint32_t* __restrict__ rs, uint32_t m, uint64_t n) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t M8 = (m + 7) & ~7u;
    const uint32_t NT = (m + 255) / 256;   // tiles of the low m digits (<= NC)
    constexpr uint32_t NB = 32 * NC;
    uint32_t* RA = lds + (size_t)w * mfma_wave_words(NC, M8);
    uint32_t* BP = RA + 16 * NB + 8;
    uint32_t* LO = RA + 8 * NB - M8;
    // zero blocks of RA (never written again)
    for (uint32_t i = 4 * lane; i < 8 * NB - M8; i += 256) *reinterpret_cast<u32x4*>(RA + i) = u32x4{0u, 0u, 0u, 0u};
    for (uint32_t i = 4 * lane; i < 8 * NB + 8; i += 256) *reinterpret_cast<u32x4*>(RA + 8 * NB + i) = u32x4{0u, 0u, 0u, 0u};
    const uint32_t t = lane & 31, h = lane >> 5;
    const uint64_t stride = (uint64_t)gridDim.x * kMfmaWaves;
    uint64_t inst = (uint64_t)blockIdx.x * kMfmaWaves + w;

    // the digits of the wave's next instance are loaded while the current
    // one is multiplied
    u32x4 xa[NC], xb[NC];
    if (inst < n) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            xa[c] = load4<VEC>(as + inst * m, 256 * c + 4 * lane, m);
            xb[c] = load4<VEC>(bs + inst * m, 256 * c + 4 * lane, m);
        }
    }
    for (; inst < n; inst += stride) {
        const uint32_t pa = top_digit<NC>(xa, lane), pb = top_digit<NC>(xb, lane);
        to_signed<NC>(xa);
        to_signed<NC>(xb);
        wave_sync();   // the previous instance's digits are out of LDS
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const uint32_t k = 256 * c + 4 * lane;
            if (k < M8) {
                *reinterpret_cast<u32x4*>(RA + 8 * NB - 4 - k) =
                    u32x4{bswap32(xa[c][3]), bswap32(xa[c][2]), bswap32(xa[c][1]), bswap32(xa[c][0])};
                *reinterpret_cast<u32x4*>(BP + 8 + k) = xb[c];
            }
        }
        if (lane < 2) *reinterpret_cast<u32x4*>(BP + 4 * lane) = u32x4{0u, 0u, 0u, 0u};
        if (lane < 7) *reinterpret_cast<u32x4*>(BP + 8 + M8 + 4 * lane) = u32x4{0u, 0u, 0u, 0u};
        const uint64_t next = inst + stride;
        if (next < n) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                xa[c] = load4<VEC>(as + next * m, 256 * c + 4 * lane, m);
                xb[c] = load4<VEC>(bs + next * m, 256 * c + 4 * lane, m);
            }
        }
        wave_sync();

        // blocks of a' / b' that may be nonzero; tiles whose column sums may be
        // nonzero (the byte product has at most 4(pa+pb)+1 columns).  Every
        // tile takes every step q (a column whose block is out of range reads
        // the zero block), so the accumulators see no branches; tiles past
        // NTa are dropped below.
        const uint32_t Pb = min(pb / 8 + 1, M8 / 8);
        const uint32_t NTa = (pa && pb) ? min(NT, (pa + pb + 2 + 255) / 256) : 0u;
        const uint32_t qend = NTa ? min(Pb, 32 * (NTa - 1) + 31) + 1 : 0;
        // column t of tile T at step q reads block p = 32T + t - q of a' (K
        // half h: bytes a'[32p + 31 - 16h - j]) at words 8(NB-1-p) + 4h =
        // colw + 8q + 256(NC-1-T)
        const uint32_t sh = (t + 1) & 3;
        const uint32_t* hsrc = BP + ((t + 16 * h + 1) >> 2);
        const uint32_t* colw = RA + 8 * (NB - 1 - 32 * (NC - 1) - t) + 4 * h;
        auto rd_b = [&](int T, uint32_t q) {
            return *reinterpret_cast<const i32x4*>(colw + 8 * q + 256 * (NC - 1 - T));
        };
        i32x4 A0 = fin_h(rd_h(hsrc, 0), sh), A1;
        i32x4 B0[NC], B1[NC];
#pragma unroll
        for (int T = 0; T < NC; T++) B0[T] = rd_b(T, 0);
        // step 0 starts the accumulators (an instance with a zero operand
        // runs it too; its tiles are dropped below), then two steps per pass
        // from step 1 (a last extra step multiplies zero blocks only: past
        // qend every column's block is below a' or past its tiles)
        i32x16 acc[NC];
        {
            const HRaw r = rd_h(hsrc, 1);
#pragma unroll
            for (int T = 0; T < NC; T++) B1[T] = rd_b(T, 1);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int T = 0; T < NC; T++) acc[T] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B0[T], i32x16{}, 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            A1 = fin_h(r, sh);
        }
        for (uint32_t q = 1; q < qend; q += 2) {
            {
                const HRaw r = rd_h(hsrc, q + 1);
#pragma unroll
                for (int T = 0; T < NC; T++) B0[T] = rd_b(T, q + 1);
                __builtin_amdgcn_sched_barrier(0);   // next step's reads go out before this step's products
#pragma unroll
                for (int T = 0; T < NC; T++)
                    acc[T] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A1, B1[T], acc[T], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                A0 = fin_h(r, sh);
            }
            {
                const HRaw r = rd_h(hsrc, q + 2);
#pragma unroll
                for (int T = 0; T < NC; T++) B1[T] = rd_b(T, q + 2);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int T = 0; T < NC; T++)
                    acc[T] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A0, B0[T], acc[T], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                A1 = fin_h(r, sh);
            }
        }
        wave_sync();   // every lane is done with the operands

        // column sums -> lo_j (LO[j]) and hi_j (BP[1 + j])
#pragma unroll
        for (int T = 0; T < NC; T++) {
            if ((uint32_t)T < NT) {
                const bool live = (uint32_t)T < NTa;
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    const uint32_t j = 256 * T + 8 * t + 2 * g + h;
                    const int64_t V = (int64_t)acc[T][4 * g] + ((int64_t)acc[T][4 * g + 1] << 8) +
                                      ((int64_t)acc[T][4 * g + 2] << 16) + ((int64_t)acc[T][4 * g + 3] << 24);
                    if (j < M8) {
                        LO[j] = live ? (uint32_t)V : 0u;
                        BP[1 + j] = live ? (uint32_t)(V >> 32) : 0u;
                    }
                }
            }
        }
        if (lane == 0) BP[0] = 0;
        wave_sync();

        // digits: d_j = lo_j + hi_(j-1) with its carry c_j in {-1, 0, 1} into
        // digit j+1; add the +1 carries, then subtract the -1 carries
        uint32_t cy1 = 0, cy2 = 0, inp = 0, inn = 0;
        uint32_t* gr = rs + inst * m;
        for (uint32_t c = 0; c < NT; c++) {
            const uint32_t k = 256 * c + 4 * lane;
            u32x4 lo = {0u, 0u, 0u, 0u}, hi = {0u, 0u, 0u, 0u};
            if (k < M8) {
                lo = *reinterpret_cast<const u32x4*>(LO + k);
                hi = *reinterpret_cast<const u32x4*>(BP + k);   // hi_(k-1) .. hi_(k+2)
            }
            u32x4 d, yp, yn;
            uint32_t cp[4], cn[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // lo + hi (hi signed): wraps up (+1) or below zero (-1)
                d[j] = lo[j] + hi[j];
                const bool neg = (int32_t)hi[j] < 0;
                cp[j] = !neg && d[j] < lo[j];
                cn[j] = neg && d[j] > lo[j];
            }
            // lane l-1's top carries (wave_shr:1; lane 0 takes the previous pass's)
            const uint32_t up2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(cp[3] | (cn[3] << 1)), 0x138, 0xF, 0xF, false);
            yp[0] = lane ? (up2 & 1u) : inp;
            yn[0] = lane ? (up2 >> 1) : inn;
#pragma unroll
            for (int j = 1; j < 4; j++) {
                yp[j] = cp[j - 1];
                yn[j] = cn[j - 1];
            }
            inp = __builtin_amdgcn_readlane(cp[3], kWave - 1);
            inn = __builtin_amdgcn_readlane(cn[3], kWave - 1);
            u32x4 s;
            bool g, p;
            uint32_t cout;
            s = lane_step<false>(d, yp, g, p);
            uint32_t e = wave_carry(g, p, cy1, cout);
            cy1 = cout;
            const u32x4 X = lane_post<false>(s, 0u, e);
            s = lane_step<true>(X, yn, g, p);
            e = wave_carry(g, p, cy2, cout);
            cy2 = cout;
            const u32x4 Y = lane_post<true>(s, 0u, e);
            if (VEC && k + 4 <= m) {
                st4<true>(gr + k, Y);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (k + j < m) gr[k + j] = Y[j];
            }
        }
    }
}

}  // namespace

bool mul_mfma_eligible(uint32_t m) { return m >= kMulMfmaMinDigits && m <= kMulMfmaMaxDigits; }

hipError_t mul_mfma_uniform(const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m, uint64_t n,
                            hipStream_t st, uint32_t num_cus) {
    if (n == 0) return hipSuccess;
    if (!mul_mfma_eligible(m)) return hipErrorInvalidValue;
    const uint32_t M8 = (m + 7) & ~7u;
    const uint32_t nc = (m + 255) / 256;
    const size_t lds = (size_t)kMfmaWaves * mfma_wave_words(nc, M8) * 4;
    const bool vec = (m % 4 == 0) && ((((uintptr_t)as) | ((uintptr_t)bs) | ((uintptr_t)rs)) & 15u) == 0;
    const void* fn;
#define BI_MFMA_FN(C) (vec ? reinterpret_cast<const void*>(&k_mul_mfma<C, true>) : reinterpret_cast<const void*>(&k_mul_mfma<C, false>))
    fn = nc <= 1 ? BI_MFMA_FN(1) : nc == 2 ? BI_MFMA_FN(2) : nc == 3 ? BI_MFMA_FN(3) : BI_MFMA_FN(4);
#undef BI_MFMA_FN
    const uint32_t per_cu = std::max<uint32_t>(1, resident_per_cu(fn, 64 * kMfmaWaves, lds));
    const uint64_t grid = std::min<uint64_t>((n + kMfmaWaves - 1) / kMfmaWaves, (uint64_t)num_cus * per_cu);
    const dim3 g((uint32_t)grid), b(64 * kMfmaWaves);
#define BI_MFMA_GO(C)                                                                  \
    do {                                                                               \
        if (vec) hipLaunchKernelGGL((k_mul_mfma<C, true>), g, b, lds, st, as, bs, rs, m, n);  \
        else hipLaunchKernelGGL((k_mul_mfma<C, false>), g, b, lds, st, as, bs, rs, m, n);     \
    } while (0)
    if (nc <= 1) BI_MFMA_GO(1);
    else if (nc == 2) BI_MFMA_GO(2);
    else if (nc == 3) BI_MFMA_GO(3);
    else BI_MFMA_GO(4);
#undef BI_MFMA_GO
    return hipGetLastError();
}

}  // namespace bi
