// bi_ntt.hip -- full products of long integers by number-theoretic transform
// over p = 2^64 - 2^32 + 1, one workgroup per product (VALU + LDS).
//
// Operands are cut into W-bit chunks.  With W = 16 (two chunks per u32
// digit) a coefficient of the chunk convolution is below min(2ma, 2mb) 2^32
// and a product of ma + mb digits needs L >= 2 (ma + mb) points; with W = 25
// the coefficients stay below min(chunks) 2^50 <= 2^63 and the product needs
// ceil(32 ma / 25) + ceil(32 mb / 25) points.  Either way every coefficient
// is below p and comes out of the transform exact.  L = 2^LOGN <= 2^14
// points fit one workgroup's LDS (8 bytes a point, 1/16 padding): ma + mb <=
// 8192 digits with 16-bit chunks, <= 12800 with 25-bit ones.
//
// Each of the L/16 lanes holds 16 points in registers.  The decimation-in-
// frequency transform (natural order in, bit-reversed order out) runs in
// passes of up to four radix-2 stages on those registers; between passes the
// points go through LDS to the next pass's groups.  The first pass reads its
// points straight from the operand digits, both forward transforms leave
// their points in the same registers, so the pointwise product needs no LDS,
// and the inverse transform (decimation in time, the stages in reverse
// order) runs the passes backwards and ends with the natural-order
// coefficients of the first pass's groups.  The digits of the product are
// then the coefficients shifted into place plus carries: block-wide carry
// look-ahead passes, as in the add.
//
// Counterpart in the reference: none.  Its GPU products are quadratic
// (cuda/ker-mul.cu.h:147-666) and its division prototype multiplies with GMP
// where a transform product was meant (prototypes/div.c:184-199, `mult`).
// The crossover against the quadratic long product is measured, not assumed.
#include <algorithm>
#include <vector>

#include "../../include/bigint.h"
#include "bi_bigmul.h"
#include "bi_ctx.h"
#include "bi_mulcore.h"
#include "bi_ntt.h"

namespace bi {

namespace {

constexpr int kNttLogTab = 14;   // twiddle tables of the 2^14-point transform
constexpr int kNttLogMin = 8;
// bi_mul_full_u32_dev picks the transform when ma * mb reaches this many
// digit products for its transform size (log2 points 12, 13, 14): the
// measured crossover with the quadratic long product (a 2^12-point transform
// product costs as much as ~1.6e6 quadratic multiply-adds, a 2^13-point one
// ~3.7e6, a 2^14-point one ~7.5e6; smaller transforms never win)
constexpr double kNttMinWork12 = 1.6e6, kNttMinWork13 = 3.7e6, kNttMinWork14 = 7.5e6;

struct NttArgs {
    const uint32_t* as;
    const uint32_t* bs;
    uint32_t* rs;
    uint64_t sa, sb, sr;   // instance strides, digits
    uint32_t ma, mb;       // operand digits
    uint32_t rlo, rlen;    // product digits [rlo, rlo + rlen) are written
    const uint64_t* wf;    // w^k, k < 2^(kNttLogTab-1), w of order 2^kNttLogTab
    const uint64_t* wi;    // w^-k
    uint64_t linv;         // L^-1 mod p
    uint64_t n;
    uint64_t* stash;       // grid * L words: the first operand's transform (large L)
    uint32_t neg;          // store (B^rlen - product) mod B^rlen (rlo == 0)
    // the second operand's transform, n * L words (instance-major, in the
    // registers' order): stored by tb_save, taken from tb_load instead of
    // being computed (the same operand, transform size and chunk width)
    uint64_t* tb_save;
    const uint64_t* tb_load;
};

// The 2^14-point transform with 25-bit chunks parks the first operand's
// transform in global memory (L2) while the second one runs: 16 more points
// a lane do not fit in its 128 registers beside the chunk extraction.  (All
// other shapes keep it in registers; parking costs more than it saves there.)
template <int LOGN, int W>
constexpr bool kStash = LOGN == 14 && W == 25;

template <int LOGN, int PASS>
struct PassGeom {
    static constexpr int NP = (LOGN + 3) / 4;
    static constexpr int R0 = LOGN - 4 * (NP - 1);
    static constexpr int T0 = PASS == 0 ? 0 : R0 + 4 * (PASS - 1);   // first stage of the pass
    static constexpr int R = PASS == 0 ? R0 : 4;                      // stages in the pass
    static constexpr int LS = LOGN - T0 - R;                          // log2 of the group stride
    static constexpr int G = 1 << R;                                  // points per group
    static constexpr int U = 16 / G;                                  // groups per lane
    static constexpr uint32_t T = 1u << (LOGN - 4);                   // lanes
};

__device__ __forceinline__ uint32_t pad(uint32_t i) { return i + (i >> 4); }

// The lane index as a value the compiler cannot see through: every LDS,
// twiddle and operand address is computed from a fresh copy where it is used,
// instead of being shared between the passes and the three transforms of a
// product and kept alive (spilled, at 2^14 points) in between.
__device__ __forceinline__ uint32_t opaque(uint32_t g) {
    asm volatile("" : "+v"(g));
    return g;
}

// point index of register k of lane g in pass PASS; lo = the group's offset
// below its stride
template <int LOGN, int PASS>
__device__ __forceinline__ uint32_t point(uint32_t g, int k) {
    using G_ = PassGeom<LOGN, PASS>;
    const uint32_t q = g + (uint32_t)(k / G_::G) * G_::T;
    const uint32_t lo = q & ((1u << G_::LS) - 1);
    const uint32_t hi = q >> G_::LS;
    return (hi << (LOGN - G_::T0)) + lo + ((uint32_t)(k % G_::G) << G_::LS);
}

// The stages of one pass on the lane's registers.  Stage t (global) pairs
// points i and i + h, h = 2^(LOGN-1-t), with twiddle w_L^((i mod h) 2^t);
// within a group (points i0 + j s) that is j and j + half with exponent
// (lo + (j mod half) s) 2^t.  Forward (DIF): (x, y) -> (x + y, (x - y) w).
// Inverse (DIT, stages in reverse): (x, y) -> (x + y w^-1, x - y w^-1).
template <int LOGN, int PASS, bool INV>
__device__ __forceinline__ void pass_stages(uint64_t* v, uint32_t g, const uint64_t* __restrict__ tw) {
    using G_ = PassGeom<LOGN, PASS>;
    constexpr int G = G_::G, R = G_::R, LS = G_::LS;
    const uint32_t gg = opaque(g);
#pragma unroll
    for (int u = 0; u < G_::U; u++) {
        const uint32_t q = gg + (uint32_t)u * G_::T;
        const uint32_t lo = q & ((1u << LS) - 1);
#pragma unroll
        for (int tt = 0; tt < R; tt++) {
            const int tau = INV ? R - 1 - tt : tt;
            const int half = G >> (tau + 1);
            const int t = G_::T0 + tau;
#pragma unroll
            for (int jj = 0; jj < half; jj++) {
                const uint32_t e = (lo + ((uint32_t)jj << LS)) << t;
                const uint64_t w = tw[(size_t)e << (kNttLogTab - LOGN)];
#pragma unroll
                for (int j0 = 0; j0 < G; j0 += 2 * half) {
                    uint64_t& x = v[u * G + j0 + jj];
                    uint64_t& y = v[u * G + j0 + jj + half];
                    if (!INV) {
                        const uint64_t a = gl::add(x, y);
                        y = gl::mul(gl::sub(x, y), w);
                        x = a;
                    } else {
                        const uint64_t b = gl::mul(y, w);
                        y = gl::sub(x, b);
                        x = gl::add(x, b);
                    }
                }
            }
        }
    }
}

template <int LOGN, int PASS>
__device__ __forceinline__ void lds_put(uint64_t* lds, const uint64_t* v, uint32_t g) {
    const uint32_t gg = opaque(g);
#pragma unroll
    for (int k = 0; k < 16; k++) lds[pad(point<LOGN, PASS>(gg, k))] = v[k];
}

template <int LOGN, int PASS>
__device__ __forceinline__ void lds_get(const uint64_t* lds, uint64_t* v, uint32_t g) {
    const uint32_t gg = opaque(g);
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = lds[pad(point<LOGN, PASS>(gg, k))];
}

template <int LOGN, int PASS>
__device__ __forceinline__ void forward_rest(uint64_t* v, uint64_t* lds, uint32_t g, const uint64_t* __restrict__ wf) {
    if constexpr (PASS < PassGeom<LOGN, 0>::NP) {
        __syncthreads();
        lds_put<LOGN, PASS - 1>(lds, v, g);
        __syncthreads();
        lds_get<LOGN, PASS>(lds, v, g);
        pass_stages<LOGN, PASS, false>(v, g, wf);
        forward_rest<LOGN, PASS + 1>(v, lds, g, wf);
    }
}

// chunk j (W bits from bit W j) of x (mx digits)
template <int W>
__device__ __forceinline__ uint64_t chunk(const uint32_t* __restrict__ x, uint32_t mx, uint32_t j) {
    if constexpr (W == 16) {
        return (j >> 1) < mx ? (uint64_t)((x[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu) : 0ull;
    } else {
        const uint32_t bit = j * W, q = bit >> 5, r = bit & 31;
        if (q >= mx) return 0ull;
        const uint64_t two = (uint64_t)x[q] | ((q + 1 < mx) ? (uint64_t)x[q + 1] << 32 : 0ull);
        return (two >> r) & ((1ull << W) - 1);
    }
}

// forward transform of the W-bit chunks of x (mx digits)
template <int LOGN, int W>
__device__ __forceinline__ void forward(uint64_t* v, const uint32_t* __restrict__ x, uint32_t mx, uint64_t* lds,
                                        uint32_t g, const uint64_t* __restrict__ wf) {
    const uint32_t gg = opaque(g);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        v[k] = chunk<W>(x, mx, point<LOGN, 0>(gg, k));
    }
    pass_stages<LOGN, 0, false>(v, g, wf);
    forward_rest<LOGN, 1>(v, lds, g, wf);
}

template <int LOGN, int PASS>
__device__ __forceinline__ void inverse_from(uint64_t* v, uint64_t* lds, uint32_t g, const uint64_t* __restrict__ wi) {
    pass_stages<LOGN, PASS, true>(v, g, wi);
    if constexpr (PASS > 0) {
        __syncthreads();
        lds_put<LOGN, PASS>(lds, v, g);
        __syncthreads();
        lds_get<LOGN, PASS - 1>(lds, v, g);
        inverse_from<LOGN, PASS - 1>(v, lds, g, wi);
    }
}

// Carry into this lane of a block-wide add whose lanes produce a carry out
// of c0 with carry-in 0 and c1 with carry-in 1: per wave by ballots, then
// across the waves through `wagg` (two words a wave).  Every lane calls it.
__device__ __forceinline__ uint32_t block_carry_in(uint32_t c0, uint32_t c1, uint32_t* wagg, uint32_t wave) {
    const bool gen = c0 != 0, prop = c1 != 0 && c0 == 0;
    uint32_t w0, w1;
    (void)wave_carry(gen, prop, 0u, w0);
    (void)wave_carry(gen, prop, 1u, w1);
    if (lane_id() == 0) {
        wagg[2 * wave] = w0;
        wagg[2 * wave + 1] = w1;
    }
    __syncthreads();
    uint32_t cw = 0;
    for (uint32_t w = 0; w < wave; w++) cw = cw ? wagg[2 * w + 1] : wagg[2 * w];
    uint32_t dummy;
    const uint32_t c = wave_carry(gen, prop, cw, dummy);
    __syncthreads();   // wagg is free again
    return c;
}

// Stores the lane's digits d0 .. d0+N-1 that fall in [rlo, rlo + rlen); with
// neg (rlo == 0) the digits of (B^rlen - value) mod B^rlen instead: digit d
// becomes ~v_d + [v_j == 0 for all j < d], the "all lower digits zero" flag
// entering the lane being a block-wide look-ahead whose every lane with a
// nonzero digit generates.
template <int N>
__device__ __forceinline__ void store_digits(uint32_t* r, uint32_t d0, uint32_t* wagg, uint32_t wave, uint32_t* out,
                                             uint32_t rlo, uint32_t rlen, uint32_t neg) {
    if (neg) {   // uniform across the workgroup
        uint32_t any = 0;
#pragma unroll
        for (int j = 0; j < N; j++) any |= r[j];
        bool z = block_carry_in(any != 0u, 1u, wagg, wave) == 0u;
#pragma unroll
        for (int j = 0; j < N; j++) {
            const uint32_t v = r[j];
            r[j] = ~v + (z ? 1u : 0u);
            z = z && v == 0u;
        }
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        const uint32_t d = d0 + j;
        if (d >= rlo && d < rlo + rlen) out[d - rlo] = r[j];
    }
}

// 16-bit chunks: lane g owns digits 8g .. 8g+7; digit d = the low word of
// c[2d] + c[2d+1] 2^16 (< 2^63) plus the high word of digit d-1 (< 2^31)
// plus the carry
__device__ __forceinline__ void digits16(const uint64_t* lds, uint32_t* wagg, uint32_t g, uint32_t wave,
                                         uint32_t* out, uint32_t rlo, uint32_t rlen, uint32_t neg) {
    uint32_t lo_[8], hi_prev[8];
    uint64_t prev = 0;
    if (g > 0) prev = lds[pad(16 * g - 2)] + (lds[pad(16 * g - 1)] << 16);
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t d = 8 * g + j;
        const uint64_t cur = lds[pad(2 * d)] + (lds[pad(2 * d + 1)] << 16);
        lo_[j] = (uint32_t)cur;
        hi_prev[j] = (uint32_t)(prev >> 32);
        prev = cur;
    }
    uint32_t c0 = 0, c1 = 1;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t s = (uint64_t)lo_[j] + hi_prev[j];
        c0 = (uint32_t)((s + c0) >> 32);
        c1 = (uint32_t)((s + c1) >> 32);
    }
    uint32_t c = block_carry_in(c0, c1, wagg, wave);
    uint32_t r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t s = (uint64_t)lo_[j] + hi_prev[j] + c;
        c = (uint32_t)(s >> 32);
        r[j] = (uint32_t)s;
    }
    store_digits<8>(r, 8 * g, wagg, wave, out, rlo, rlen, neg);
}

// 25-bit chunks: lane g owns digits 13g .. 13g+12.  S_e = the sum of
// c_j 2^(25j - 32e) over the one or two chunks j that start inside digit e
// (c_j < 2^63, so S_e < 2^95: three words); the product is
// sum_e S_e B^e = low words + middle words B + high words B^2, added in two
// carry look-ahead passes (each with one-bit carries).  The S_e are
// recomputed from LDS in each pass rather than held in registers.
template <int LOGN>
__device__ __forceinline__ void s_words(const uint64_t* lds, int32_t e, uint32_t& w0, uint32_t& w1, uint32_t& w2) {
    constexpr uint32_t L = 1u << LOGN;
    w0 = w1 = w2 = 0;
    if (e < 0) return;
    const uint32_t j0 = (32u * (uint32_t)e + 24) / 25;   // first chunk starting in digit e
    const uint32_t j1 = (32u * (uint32_t)e + 31) / 25;   // last one (j0 or j0 + 1)
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const uint32_t j = j0 + u;
        if (j <= j1 && j < L) {
            const uint64_t c = lds[pad(j)];
            const uint32_t sh = 25 * j - 32 * (uint32_t)e;   // 0 .. 31
            const uint64_t lo = c << sh;
            const uint32_t hi = sh ? (uint32_t)(c >> (64 - sh)) : 0u;
            const uint64_t t0 = (uint64_t)w0 + (uint32_t)lo;
            const uint64_t t1 = (uint64_t)w1 + (uint32_t)(lo >> 32) + (t0 >> 32);
            w0 = (uint32_t)t0;
            w1 = (uint32_t)t1;
            w2 += hi + (uint32_t)(t1 >> 32);
        }
    }
}

template <int LOGN>
__device__ __forceinline__ void digits25(const uint64_t* lds, uint32_t* wagg, uint32_t g, uint32_t wave,
                                         uint32_t* out, uint32_t rlo, uint32_t rlen, uint32_t neg) {
    constexpr int DPL = 13;
    const int32_t d0 = (int32_t)(DPL * g);
    uint32_t r[DPL];
    uint32_t x0, x1, x2;
    {   // low words + middle words one digit up
        uint32_t c0 = 0, c1 = 1, m_prev;
        s_words<LOGN>(lds, d0 - 1, x0, m_prev, x2);
#pragma unroll
        for (int k = 0; k < DPL; k++) {
            s_words<LOGN>(lds, d0 + k, x0, x1, x2);
            r[k] = x0;
            const uint64_t a = (uint64_t)x0 + m_prev;
            m_prev = x1;
            c0 = (uint32_t)((a + c0) >> 32);
            c1 = (uint32_t)((a + c1) >> 32);
        }
        uint32_t c = block_carry_in(c0, c1, wagg, wave);
        s_words<LOGN>(lds, d0 - 1, x0, m_prev, x2);
#pragma unroll
        for (int k = 0; k < DPL; k++) {
            s_words<LOGN>(lds, d0 + k, x0, x1, x2);
            const uint64_t a = (uint64_t)r[k] + m_prev + c;
            m_prev = x1;
            r[k] = (uint32_t)a;
            c = (uint32_t)(a >> 32);
        }
    }
    {   // + high words two digits up
        uint32_t c0 = 0, c1 = 1;
#pragma unroll
        for (int k = 0; k < DPL; k++) {
            s_words<LOGN>(lds, d0 + k - 2, x0, x1, x2);
            const uint64_t a = (uint64_t)r[k] + x2;
            c0 = (uint32_t)((a + c0) >> 32);
            c1 = (uint32_t)((a + c1) >> 32);
        }
        uint32_t c = block_carry_in(c0, c1, wagg, wave);
#pragma unroll
        for (int k = 0; k < DPL; k++) {
            s_words<LOGN>(lds, d0 + k - 2, x0, x1, x2);
            const uint64_t a = (uint64_t)r[k] + x2 + c;
            c = (uint32_t)(a >> 32);
            r[k] = (uint32_t)a;
        }
    }
    store_digits<DPL>(r, (uint32_t)d0, wagg, wave, out, rlo, rlen, neg);
}

// At least four waves a SIMD (at most 128 registers a lane, a few spilled):
// with 145-150 registers the 2^13-point kernel ran one 512-lane workgroup a
// CU, two waves a SIMD, and the latency of its dependent 64-bit chains went
// unhidden; two workgroups a CU took the division batch's transforms from
// 17.9 to about 14.5 ms (16384 divisions of 8192 / 4096 digits: 37.7 -> 34.4
// ms).  Cutting the multiply-adds instead (radix-16 passes with
// power-of-two twiddles) had changed nothing.
template <int LOGN, int W>
__global__ void __launch_bounds__(1 << (LOGN - 4), 4) k_ntt_mul(NttArgs A) {
    constexpr int L = 1 << LOGN;
    constexpr int NP = PassGeom<LOGN, 0>::NP;
    extern __shared__ uint64_t lds[];   // pad(L) points, then 2 words per wave
    uint32_t* wagg = reinterpret_cast<uint32_t*>(lds + L + L / 16);
    const uint32_t wave = threadIdx.x / kWave;
    uint64_t acc[16], v[16];
    for (uint64_t inst = blockIdx.x; inst < A.n; inst += gridDim.x) {
        const uint32_t g = opaque(threadIdx.x);
        forward<LOGN, W>(acc, A.as + inst * A.sa, A.ma, lds, g, A.wf);
        uint64_t* park = A.stash + (size_t)blockIdx.x * L + g;
        if constexpr (kStash<LOGN, W>) {
#pragma unroll
            for (int k = 0; k < 16; k++) park[(size_t)k * (L / 16)] = acc[k];
            forward<LOGN, W>(v, A.bs + inst * A.sb, A.mb, lds, g, A.wf);
#pragma unroll
            for (int k = 0; k < 16; k++) acc[k] = park[(size_t)k * (L / 16)];
        } else if (A.tb_load) {
            const uint64_t* t = A.tb_load + inst * L + g;
#pragma unroll
            for (int k = 0; k < 16; k++) v[k] = t[(size_t)k * (L / 16)];
        } else {
            forward<LOGN, W>(v, A.bs + inst * A.sb, A.mb, lds, g, A.wf);
            if (A.tb_save) {
                uint64_t* t = A.tb_save + inst * L + g;
#pragma unroll
                for (int k = 0; k < 16; k++) t[(size_t)k * (L / 16)] = v[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 16; k++) acc[k] = gl::mul(acc[k], v[k]);
        inverse_from<LOGN, NP - 1>(acc, lds, g, A.wi);
        // natural-order coefficients c_i = acc / L, into LDS
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) lds[pad(point<LOGN, 0>(g, k))] = gl::mul(acc[k], A.linv);
        __syncthreads();
        uint32_t* out = A.rs + inst * A.sr;
        if constexpr (W == 16) digits16(lds, wagg, g, wave, out, A.rlo, A.rlen, A.neg);
        else digits25<LOGN>(lds, wagg, g, wave, out, A.rlo, A.rlen, A.neg);
    }
}

template <int LOGN>
size_t ntt_lds_bytes() {
    constexpr int L = 1 << LOGN;
    constexpr uint32_t T = 1u << (LOGN - 4);
    return (size_t)(L + L / 16) * 8 + ((T + 63) / 64) * 8;
}

template <int LOGN>
uint64_t ntt_grid(uint32_t num_cus, uint64_t n) {
    const uint32_t T = 1u << (LOGN - 4);
    const uint32_t lds = (uint32_t)ntt_lds_bytes<LOGN>();
    const uint32_t per_cu = std::max<uint32_t>(1, std::min<uint32_t>((160u * 1024) / lds, 2048 / T));
    return std::min<uint64_t>(n, (uint64_t)num_cus * per_cu);
}

// stash words the transform (2^lg points, w-bit chunks) needs for n products
uint64_t ntt_stash_words(int lg, int w, uint32_t num_cus, uint64_t n) {
    return (lg == 14 && w == 25) ? ntt_grid<14>(num_cus, n) << 14 : 0;
}

template <int LOGN, int W>
void ntt_attr() {
    set_max_lds(reinterpret_cast<const void*>(&k_ntt_mul<LOGN, W>), 160 * 1024);
}

// the LDS limit of every transform kernel, set once (outside any capture)
void ntt_attrs() {
    static const bool done = [] {   // thread-safe static initialisation
        ntt_attr<8, 16>();
        ntt_attr<9, 16>();
        ntt_attr<10, 16>();
        ntt_attr<11, 16>();
        ntt_attr<12, 16>();
        ntt_attr<13, 16>();
        ntt_attr<14, 16>();
        ntt_attr<12, 25>();
        ntt_attr<13, 25>();
        ntt_attr<14, 25>();
        return true;
    }();
    (void)done;
}

template <int LOGN, int W>
hipError_t launch_ntt(NttArgs A, uint32_t num_cus, hipStream_t st) {
    const size_t lds = ntt_lds_bytes<LOGN>();
    const uint64_t grid = ntt_grid<LOGN>(num_cus, A.n);
    hipLaunchKernelGGL((k_ntt_mul<LOGN, W>), dim3((uint32_t)grid), dim3(1u << (LOGN - 4)), lds, st, A);
    return hipGetLastError();
}

// Twiddle tables of the context's device, uploaded once (scratch slot 10).
int ntt_tables(bi_ctx* ctx, const uint64_t** wf, const uint64_t** wi) {
    constexpr size_t H = (size_t)1 << (kNttLogTab - 1);
    const bool fresh = ctx->scratch[10] == nullptr;
    void* p;
    int rc = scratch(ctx, 10, 2 * H * 8, &p);
    if (rc) return rc;
    if (fresh) {
        std::vector<uint64_t> h(2 * H);
        const uint64_t w = gl::root_of_unity(kNttLogTab), winv = gl::inverse(w);
        uint64_t a = 1, b = 1;
        for (size_t k = 0; k < H; k++) {
            h[k] = a;
            h[H + k] = b;
            a = gl::mul(a, w);
            b = gl::mul(b, winv);
        }
        hipError_t e = hipMemcpyAsync(p, h.data(), 2 * H * 8, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(ctx->scratch[10]);
            ctx->scratch[10] = nullptr;
            ctx->scratch_bytes[10] = 0;
            return hip_fail(ctx, e);
        }
    }
    *wf = (const uint64_t*)p;
    *wi = (const uint64_t*)p + H;
    return BI_OK;
}

}  // namespace

// The smallest transform (log2 points, chunk bits) that holds an ma x mb
// full product: 16-bit chunks need 2 (ma + mb) points; 25-bit chunks need
// ceil(32 ma / 25) + ceil(32 mb / 25) points (their coefficients stay below
// min(chunks) 2^50 <= 2^63 < p for up to 2^14 points) and a longer digit
// phase, so they are taken only where they need fewer points (from 2^12).
// lg = -1 when no transform of up to 2^14 points holds the product.
NttPlan ntt_plan(uint32_t ma, uint32_t mb) {
    auto lg_for = [](uint64_t points) {
        int lg = kNttLogMin;
        while ((1ull << lg) < points) lg++;
        return lg;
    };
    const int l16 = lg_for(2ull * ((uint64_t)ma + mb));
    const int l25 = std::max(12, lg_for(((uint64_t)ma * 32 + 24) / 25 + ((uint64_t)mb * 32 + 24) / 25));
    if (l16 <= kNttLogTab && l16 <= l25) return {l16, 16};
    if (l25 <= kNttLogTab) return {l25, 25};
    return {-1, 0};
}

// Twiddle tables and the stash for transforms of up to 2^14 points over n
// products: everything ntt_launch allocates or uploads, done ahead of it (so
// that the launch itself can be captured in a graph).
int ntt_prepare(bi_ctx* ctx, uint64_t n, const uint64_t** wf, const uint64_t** wi, uint64_t** stash) {
    ntt_attrs();
    int rc = ntt_tables(ctx, wf, wi);
    if (rc) return rc;
    void* p;
    if ((rc = scratch(ctx, 11, ntt_stash_words(14, 25, (uint32_t)ctx->num_cus, n) * 8, &p))) return rc;
    *stash = (uint64_t*)p;
    return BI_OK;
}

// digits [rlo, rlo + rlen) of a*b (neg: of -a*b mod B^rlen, rlo == 0) for n
// pairs, by transform; ntt_prepare(ctx, n) must have run.  Allocates nothing.
hipError_t ntt_launch(bi_ctx* ctx, uint64_t n, uint32_t ma, uint32_t mb, const uint32_t* as, uint64_t sa,
                      const uint32_t* bs, uint64_t sb, uint32_t* rs, uint64_t sr, uint32_t rlo, uint32_t rlen,
                      bool neg, hipStream_t st, bool cyclic, uint64_t* tb_save, const uint64_t* tb_load) {
    // cyclic: the chunk convolution is allowed to wrap mod x^(2^13) - 1; the
    // digit pass then covers 13 digits on each of its 512 lanes
    const NttPlan pl = cyclic ? NttPlan{13, 25} : ntt_plan(ma, mb);
    const int lg = pl.lg;
    const uint64_t cover = cyclic ? 13ull * 512 : (uint64_t)ma + mb;
    if (cyclic && (ma > kCyclicDigits || mb > kCyclicDigits || neg)) return hipErrorInvalidValue;
    if (lg < 0 || (uint64_t)rlo + rlen > cover || (neg && rlo != 0)) return hipErrorInvalidValue;
    const uint64_t need = ntt_stash_words(lg, pl.w, (uint32_t)ctx->num_cus, n) * 8;
    if (!ctx->scratch[10] || (need && ctx->scratch_bytes[11] < need)) return hipErrorInvalidValue;
    constexpr size_t H = (size_t)1 << (kNttLogTab - 1);
    NttArgs A;
    A.as = as;
    A.bs = bs;
    A.rs = rs;
    A.sa = sa;
    A.sb = sb;
    A.sr = sr;
    A.ma = ma;
    A.mb = mb;
    A.rlo = rlo;
    A.rlen = rlen;
    A.n = n;
    A.wf = (const uint64_t*)ctx->scratch[10];
    A.wi = A.wf + H;
    A.linv = gl::inverse(1ull << lg);
    A.stash = (uint64_t*)ctx->scratch[11];
    A.neg = neg ? 1u : 0u;
    // a kept transform is only valid for the 2^13-point, 25-bit-chunk shape
    // (the second operand's points in registers, no stash)
    if ((tb_save || tb_load) && !(lg == 13 && pl.w == 25)) return hipErrorInvalidValue;
    A.tb_save = tb_save;
    A.tb_load = tb_load;
    const uint32_t cus = (uint32_t)ctx->num_cus;
    if (pl.w == 25) {
        switch (lg) {
            case 12: return launch_ntt<12, 25>(A, cus, st);
            case 13: return launch_ntt<13, 25>(A, cus, st);
            default: return launch_ntt<14, 25>(A, cus, st);
        }
    }
    switch (lg) {
        case 8: return launch_ntt<8, 16>(A, cus, st);
        case 9: return launch_ntt<9, 16>(A, cus, st);
        case 10: return launch_ntt<10, 16>(A, cus, st);
        case 11: return launch_ntt<11, 16>(A, cus, st);
        case 12: return launch_ntt<12, 16>(A, cus, st);
        case 13: return launch_ntt<13, 16>(A, cus, st);
        default: return launch_ntt<14, 16>(A, cus, st);
    }
}

}  // namespace bi

using namespace bi;

extern "C" {

int bi_mul_full_u32_dev(bi_ctx* ctx, uint64_t n, uint32_t ma, uint32_t mb, const uint32_t* as, const uint32_t* bs,
                        uint32_t* rs, int algo) {
    if (!ctx || ma == 0 || mb == 0 || algo < 0 || algo > 2) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    if (!as || !bs || !rs) return BI_ERR_ARG;
    if (n > (1ull << 29)) return BI_ERR_SHAPE;   // the quadratic product's grid is n times its chunks
    int rc = set_device(ctx);
    if (rc) return rc;
    const uint32_t mr = ma + mb;
    const NttPlan pl = ntt_plan(ma, mb);
    const bool fits_ntt = pl.lg >= 0;
    // the schoolbook long product stages both operands in one workgroup's LDS
    const bool fits_quad = (size_t)(((ma + 7) & ~7u) + 12 + kWinPad + ((mb + 7) & ~7u) + 16) * 4 <= 160 * 1024;
    bool use_ntt;
    if (algo == 1) use_ntt = false;
    else if (algo == 2) use_ntt = true;
    else {
        const double work = (double)ma * mb;
        use_ntt = (pl.lg == 12 && work >= kNttMinWork12) || (pl.lg == 13 && work >= kNttMinWork13) ||
                  (pl.lg == 14 && work >= kNttMinWork14);
    }
    if (use_ntt && !fits_ntt) return BI_ERR_SHAPE;
    if (!use_ntt && !fits_quad) return BI_ERR_SHAPE;
    if (use_ntt) {
        const uint64_t *wf, *wi;
        uint64_t* stash;
        if ((rc = ntt_prepare(ctx, n, &wf, &wi, &stash))) return rc;
        const hipError_t e = ntt_launch(ctx, n, ma, mb, as, ma, bs, mb, rs, mr, 0, mr, false, ctx->stream);
        return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
    }
    BigEnv env = {};
    env.num_cus = (uint32_t)ctx->num_cus;
    env.st = ctx->stream;
    env.rows = (int)ctx->mul_cfg.rows;
    if ((rc = mul_snaps(ctx, &env.snaps, &env.snap_words))) return rc;
    void* cs;
    env.cs_words = bigmul_scratch_words(ma, mb, n);
    if ((rc = scratch(ctx, 3, env.cs_words * 4, &cs))) return rc;
    env.cs = (uint32_t*)cs;
    const hipError_t e = bigmul(Win{as, ma, 0, ma}, Win{bs, mb, 0, mb}, rs, mr, 0, mr, n, env);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

}  // extern "C"
