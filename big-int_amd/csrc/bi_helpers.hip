// bi_helpers.hip -- compare / precision / shift / multiply-by-digit /
// divide-by-digit helpers over [n][m] batches.
//
// Reference semantics (prototypes/div.c and futhark/helper.fut):
//   lt/eq/ez  div.c:49-69, helper.fut:83-91 (most significant difference decides)
//   findk     div.c:92-96 (index of the top nonzero digit, 0 for zero); prec = findk+1
//   shift     div.c:105-117, helper.fut:96-98 (s >= 0: toward more significant)
//   multd     div.c:202-218 ((u*d) mod B^m); quod div.c:165-181 (floor(u/d))
#include "bi_div1.h"
#include "bi_kernels.h"

namespace bi {

// ---------------------------------------------------------------------------
// Top-down digit scan: one wave per instance looks at the top 64 digits (one
// a lane), then walks 256-digit tiles further down, and stops at the first
// tile holding a digit with pred(a, b) (a != b for compare, a != 0 for
// precision).  On random data the top digits decide, so the scan reads
// 256 bytes per operand and instance whatever m is.
//   mode 0: out_i8[i] = cmp(a, b) in {-1, 0, 1}
//   mode 1: out_u8[i] = (a < b)
//   mode 2: out_u32[i] = index of the top nonzero digit of a, +1 (0 if a == 0)
template <int MODE>
__global__ void __launch_bounds__(256) k_topscan(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                 uint32_t m, uint64_t n, void* __restrict__ out) {
    const uint32_t lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (i >= n) return;
    const uint32_t* a = as + i * m;
    const uint32_t* b = (MODE == 2) ? nullptr : bs + i * m;
    uint32_t found = 0;   // top matching index + 1
    const uint32_t top = m > kWave ? m - kWave : 0;   // digits [top, m) first
    {
        const uint32_t k = top + lane;
        uint32_t best = 0;
        if (k < m) {
            const bool hit = (MODE == 2) ? (a[k] != 0u) : (a[k] != b[k]);
            best = hit ? k + 1 : 0;
        }
        found = readfirst(wave_max_u32(best));
    }
    for (int64_t tb = top ? (int64_t)((top - 1) & ~255u) : -1; tb >= 0 && !found; tb -= 256) {
        uint32_t best = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = (uint32_t)tb + 4 * lane + j;
            if (k < top) {
                const bool hit = (MODE == 2) ? (a[k] != 0u) : (a[k] != b[k]);
                best = hit ? k + 1 : best;
            }
        }
        found = readfirst(wave_max_u32(best));
    }
    if (lane != 0) return;
    if (MODE == 2) {
        static_cast<uint32_t*>(out)[i] = found;
    } else {
        int r = 0;
        if (found) r = a[found - 1] < b[found - 1] ? -1 : 1;
        if (MODE == 0) static_cast<int8_t*>(out)[i] = (int8_t)r;
        else static_cast<uint8_t*>(out)[i] = (uint8_t)(r < 0);
    }
}

hipError_t topscan(int mode, const uint32_t* as, const uint32_t* bs, uint32_t m, uint64_t n, void* out,
                   hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid((uint32_t)((n + 3) / 4)), block(256);
    if (mode == 0) hipLaunchKernelGGL(k_topscan<0>, grid, block, 0, st, as, bs, m, n, out);
    else if (mode == 1) hipLaunchKernelGGL(k_topscan<1>, grid, block, 0, st, as, bs, m, n, out);
    else hipLaunchKernelGGL(k_topscan<2>, grid, block, 0, st, as, bs, m, n, out);
    return hipGetLastError();
}

// prec = max(sig, 1): the reference's findk+1 (zero has precision 1); in
// u64 limbs (U64) the significant u32 digits d give ceil(d / 2) limbs
template <bool U64>
__global__ void k_prec_fix(uint32_t* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t d = p[i] ? p[i] : 1u;
        p[i] = U64 ? (d + 1) >> 1 : d;
    }
}

hipError_t prec_ref(uint32_t* p, uint64_t n, hipStream_t st, bool u64_limbs) {
    if (n == 0) return hipSuccess;
    const dim3 grid((uint32_t)((n + 255) / 256)), block(256);
    if (u64_limbs) hipLaunchKernelGGL(k_prec_fix<true>, grid, block, 0, st, p, n);
    else hipLaunchKernelGGL(k_prec_fix<false>, grid, block, 0, st, p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Digit shift: rs[i][k] = as[i][k - s_i] inside [0, m), zero elsewhere.  One
// thread per output group of 4 digits; s_i is `s` or sarr[i] when given.
// VEC (m % 4 == 0, rs 16-byte aligned): the group is one 16-byte
// nontemporal store (the source digits, shifted, are four coalesced loads).
template <bool VEC>
__global__ void __launch_bounds__(256) k_shift(const uint32_t* __restrict__ as, uint32_t* __restrict__ rs, uint32_t m,
                                               uint64_t n, int32_t s, const int32_t* __restrict__ sarr) {
    const uint64_t groups_per = (m + 3) / 4;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups_per * n) return;
    const uint64_t i = g / groups_per;
    const uint32_t k0 = (uint32_t)(g - i * groups_per) * 4;
    const int64_t sh = sarr ? sarr[i] : s;
    const uint32_t* a = as + i * m;
    uint32_t* r = rs + i * m;
    u32x4 v;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int64_t k = (int64_t)k0 + j;
        const int64_t src = k - sh;
        v[j] = (k < m && src >= 0 && src < (int64_t)m) ? a[src] : 0u;
    }
    if (VEC) {
        st4<true>(r + k0, v);
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (k0 + j < m) r[k0 + j] = v[j];
    }
}

hipError_t shift(const uint32_t* as, uint32_t* rs, uint32_t m, uint64_t n, int32_t s, const int32_t* sarr,
                 hipStream_t st) {
    const uint64_t total = ((uint64_t)(m + 3) / 4) * n;
    if (total == 0) return hipSuccess;
    const bool vec = (m % 4) == 0 && ((uintptr_t)rs & 15) == 0;
    const dim3 grid((uint32_t)((total + 255) / 256)), block(256);
    if (vec) hipLaunchKernelGGL(k_shift<true>, grid, block, 0, st, as, rs, m, n, s, sarr);
    else hipLaunchKernelGGL(k_shift<false>, grid, block, 0, st, as, rs, m, n, s, sarr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// (a * d_i) mod B^m.  One wave per instance-span, 4 digits per lane: digit k
// receives lo(a_k d) + hi(a_{k-1} d); the 33-bit sums are resolved by the
// wave carry look-ahead.  The high word of the digit below a lane's first
// digit comes from the neighbouring lane (or the previous tile).  VEC (m % 4
// == 0, both pointers 16-byte aligned): 16-byte nontemporal loads and
// stores, the next tile's load issued before the current tile is resolved.
template <bool VEC>
__global__ void __launch_bounds__(256) k_muld(const uint32_t* __restrict__ as, const uint32_t* __restrict__ ds,
                                              uint32_t* __restrict__ rs, uint32_t m, uint64_t n) {
    const uint32_t lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (i >= n) return;
    const uint32_t* a = as + i * m;
    uint32_t* r = rs + i * m;
    const uint32_t d = ds[i];
    uint32_t carry = 0, hi_prev_tile = 0;
    u32x4 nxt = {0u, 0u, 0u, 0u};
    if (VEC && 4 * lane < m) nxt = ld4<true>(a + 4 * lane);
    for (uint32_t tb = 0; tb < m; tb += 256) {
        uint32_t lo[4], hi[4];
        u32x4 cur = nxt;
        if (VEC) {
            const uint32_t kn = tb + 256 + 4 * lane;
            nxt = kn < m ? ld4<true>(a + kn) : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = tb + 4 * lane + j;
            const uint64_t p = (k < m) ? (uint64_t)(VEC ? cur[j] : a[k]) * d : 0ull;
            lo[j] = (uint32_t)p;
            hi[j] = (uint32_t)(p >> 32);
        }
        uint32_t below = __shfl_up(hi[3], 1, kWave);
        if (lane == 0) below = hi_prev_tile;
        hi_prev_tile = __shfl(hi[3], 63, kWave);
        const uint32_t y[4] = {below, hi[0], hi[1], hi[2]};
        uint32_t s[4], c = 0, all = 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t t = (uint64_t)lo[j] + y[j] + c;
            s[j] = (uint32_t)t;
            c = (uint32_t)(t >> 32);
            all &= s[j];
        }
        uint32_t cout;
        uint32_t e = wave_carry(c != 0, all == 0xFFFFFFFFu, carry, cout);
        carry = cout;
        u32x4 o;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            o[j] = s[j] + e;
            e &= (uint32_t)(s[j] == 0xFFFFFFFFu);
        }
        const uint32_t k = tb + 4 * lane;
        if (VEC) {
            if (k < m) st4<true>(r + k, o);
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (k + j < m) r[k + j] = o[j];
        }
    }
}

hipError_t muld(const uint32_t* as, const uint32_t* ds, uint32_t* rs, uint32_t m, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const bool vec = (m % 4) == 0 && (((uintptr_t)as | (uintptr_t)rs) & 15) == 0;
    const dim3 grid((uint32_t)((n + 3) / 4)), block(256);
    if (vec) hipLaunchKernelGGL(k_muld<true>, grid, block, 0, st, as, ds, rs, m, n);
    else hipLaunchKernelGGL(k_muld<false>, grid, block, 0, st, as, ds, rs, m, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// floor(a / d_i) and a mod d_i: schoolbook long division from the top digit,
// one lane per instance (the digits of one quotient depend on each other).
// L lanes per instance (8 to 64: one to eight instances a wave), tiles
// of L G digits from the top (G consecutive digits a lane).  The remainder
// entering digit k is the value of the digits above k (mod d), so the L
// lanes of a tile get theirs from a suffix scan: lane j first reduces its G
// digits mod d (Horner, keeping the quotient digits of that division), then
// combines with the lanes above by V = V_above * B^(G*len) + V (mod d) --
// the powers B^(G*2^j) mod d are the same for every lane of an instance --
// and lane L-1 starts from the remainder carried in from the tile above.
// With x the remainder entering a lane's chunk X = q0 dn + r0 (q0, r0 from
// the first division), x B^G + X = dn (x Qg + q0) + (x Rg + r0) for
// B^G = Qg dn + Rg, so the lane's quotient digits are x Qg + q0 + c with
// c = floor((x Rg + r0) / dn): one two-by-one division and a G-digit
// multiply-add instead of a second division of every digit.  Instance setup
// (the reciprocal, Qg and the powers) is the same instruction stream for all
// the instances a wave holds, so several instances a wave (and wider lanes,
// scanning once per L G digits) do less work a digit.  Every division step
// is a two-digit by one-digit division with a precomputed reciprocal by the
// normalised divisor (bi_div1.h).  Division by zero is defined as quotient
// 0, remainder 0.  VEC (m % 4 == 0, both pointers 16-byte aligned): 16-byte
// nontemporal loads and stores, the next tile's load issued before the
// current tile is divided.
template <bool VEC, int G, int L>
__global__ void __launch_bounds__(256) k_quod(const uint32_t* __restrict__ as, const uint32_t* __restrict__ ds,
                                              uint32_t* __restrict__ rs, uint32_t* __restrict__ rem, uint32_t m,
                                              uint64_t n) {
    constexpr uint32_t T = L * G;   // digits a tile: G consecutive digits a lane
    constexpr int LOGL = L == 64 ? 6 : L == 32 ? 5 : L == 16 ? 4 : 3;
    static_assert(L == 64 || L == 32 || L == 16 || L == 8, "lanes per instance");
    static_assert(G % 4 == 0, "16-byte groups of digits");
    const uint32_t lane = lane_id() % L;
    const uint64_t i = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave) * (kWave / L) +
                       lane_id() / L;
    if (i >= n) return;
    const uint32_t* a = as + i * m;
    uint32_t* r = rs + i * m;
    const uint32_t d = ds[i];
    if (d == 0) {
        for (uint32_t k = lane; k < m; k += L) r[k] = 0;
        if (rem && lane == 0) rem[i] = 0;
        return;
    }
    // The division runs on N = a * 2^s by dn = d * 2^s (bi_div1.h): the same
    // quotient, residues mod dn, no shifts in the steps; the remainder is the
    // last residue >> s.  N has one more digit than a, at position m: inside
    // the top tile (from the zero digits above m) unless m % T == 0, when it
    // is the carry into the top tile.
    const Div1 D = div1_make(d);
    const uint32_t s = D.s;
    // Qg = floor(B^G / dn) (G digits), Rg = B^G mod dn
    uint32_t Qg[G], Rg = 1u;
#pragma unroll
    for (int j = G - 1; j >= 0; j--) div21(Rg, 0u, D.dn, D.v, Qg[j], Rg);
    uint32_t P[LOGL];   // P[j] = B^(G * 2^j) mod dn
    P[0] = Rg;
#pragma unroll
    for (int j = 1; j < LOGL; j++) P[j] = mulmod_n(D, P[j - 1], P[j - 1]);
    uint32_t carry = (m % T == 0 && s) ? a[m - 1] >> (32 - s) : 0u;
    const int64_t tb0 = (int64_t)((m - 1) / T * T);
    u32x4 nxt[G / 4];
#pragma unroll
    for (int t = 0; t < G / 4; t++) {
        nxt[t] = u32x4{0u, 0u, 0u, 0u};
        if (VEC && (uint32_t)tb0 + G * lane + 4 * t < m) nxt[t] = ld4<true>(a + (uint32_t)tb0 + G * lane + 4 * t);
    }
    for (int64_t tb = tb0; tb >= 0; tb -= T) {
        const uint32_t k0 = (uint32_t)tb + G * lane;
        uint32_t u[G];
        if (VEC) {
#pragma unroll
            for (int t = 0; t < G / 4; t++) {
                const u32x4 cur = nxt[t];
                // the tile below lies wholly below tb <= m
                if (tb >= (int64_t)T) nxt[t] = ld4<true>(a + k0 - T + 4 * t);
#pragma unroll
                for (int j = 0; j < 4; j++) u[4 * t + j] = (k0 + 4 * t + j < m) ? cur[j] : 0u;
            }
        } else {
#pragma unroll
            for (int j = 0; j < G; j++) u[j] = (k0 + j < m) ? a[k0 + j] : 0u;
        }
        // normalised digits: digit k of N takes the top s bits of digit k-1
        uint32_t below = __shfl_up(u[G - 1], 1, L);
        if (lane == 0) below = tb > 0 ? a[tb - 1] : 0u;
        uint32_t un[G];
        un[0] = norm_digit(u[0], below, s);
#pragma unroll
        for (int j = 1; j < G; j++) un[j] = norm_digit(u[j], u[j - 1], s);
        // the lane's chunk divided with remainder 0: quotient digits q0, residue b
        uint32_t q0[G], b = 0;
#pragma unroll
        for (int j = G - 1; j >= 0; j--) div21(b, un[j], D.dn, D.v, q0[j], b);
        // residues of the lanes' chunks, combined upward: after the scan lane
        // j holds the residue of its chunk and everything above it
        uint32_t V = b;
        if (lane == L - 1) V = addmod_n(D, mulmod_n(D, carry, P[0]), b);
#pragma unroll
        for (int j = 0; j < LOGL; j++) {
            const uint32_t o = __shfl_down(V, 1 << j, L);
            if (lane + (1u << j) < L) V = addmod_n(D, mulmod_n(D, o, P[j]), V);
        }
        const uint32_t above = __shfl_down(V, 1, L);
        const uint32_t x = (lane == L - 1) ? carry : above;
        // quotient digits x Qg + q0 + c, c = floor((x Rg + b) / dn)
        const uint64_t t = (uint64_t)x * Rg + b;   // < dn (dn + 1) <= 2^64
        uint32_t c, rdummy;
        div21((uint32_t)(t >> 32), (uint32_t)t, D.dn, D.v, c, rdummy);
        uint32_t qs[G];
        uint64_t acc = c;
#pragma unroll
        for (int j = 0; j < G; j++) {
            acc += (uint64_t)x * Qg[j] + q0[j];
            qs[j] = (uint32_t)acc;
            acc >>= 32;
        }
        if (VEC) {
#pragma unroll
            for (int t4 = 0; t4 < G / 4; t4++)
                if (k0 + 4 * t4 < m)
                    st4<true>(r + k0 + 4 * t4, u32x4{qs[4 * t4], qs[4 * t4 + 1], qs[4 * t4 + 2], qs[4 * t4 + 3]});
        } else {
#pragma unroll
            for (int j = 0; j < G; j++)
                if (k0 + j < m) r[k0 + j] = qs[j];
        }
        carry = __shfl(V, 0, L);   // lane 0: the whole tile and everything above, mod dn
    }
    if (rem && lane == 0) rem[i] = carry >> s;
}

hipError_t quod(const uint32_t* as, const uint32_t* ds, uint32_t* rs, uint32_t* rem, uint32_t m, uint64_t n,
                hipStream_t st) {
    if (n == 0) return hipSuccess;
    const bool vec = (m % 4) == 0 && (((uintptr_t)as | (uintptr_t)rs) & 15) == 0;
    // digits a lane G and lanes an instance L: the residue scan runs once a
    // tile of L G digits and the instance setup once for the 64 / L
    // instances of a wave.  Swept on one MI355X (2^26-digit batches, us):
    //   m       64   128   256   512  1024  2048
    //   G4 L8   129   111   110   112   145   152
    //   G4 L16  184   139   115   107   131   129
    //   G4 L32  377   191   140   119   113   112
    // (the former G 4/8/16 with 64 lanes: 741, 388, 200, 133, 117, 115;
    // 16 digits a lane lose everywhere: their Qg is 16 dependent divisions)
#define BI_QUOD(G, L)                                                                                        \
    do {                                                                                                     \
        const dim3 grid((uint32_t)((n + 4 * (64 / L) - 1) / (4 * (64 / L)))), block(256);                  \
        if (vec) hipLaunchKernelGGL((k_quod<true, G, L>), grid, block, 0, st, as, ds, rs, rem, m, n);        \
        else hipLaunchKernelGGL((k_quod<false, G, L>), grid, block, 0, st, as, ds, rs, rem, m, n);           \
    } while (0)
    if (m > 512) BI_QUOD(4, 32);
    else if (m > 256) BI_QUOD(4, 16);
    else BI_QUOD(4, 8);
#undef BI_QUOD
    return hipGetLastError();
}

}  // namespace bi
