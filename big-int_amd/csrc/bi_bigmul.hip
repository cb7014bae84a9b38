// bi_bigmul.hip -- batched products of long integers and windowed
// add/subtract; the building blocks of the division pipeline.
//
// rectangle (bigmul): out[k] = digit k of a_i * b_i for k in [lo, lo+len).
//   pass 1 (k_bigconv): each workgroup stages BOTH operands of its instance in
//     LDS and runs a chunk of the instance's jobs (rectangular schedule: every
//     job walks the same number of row blocks); the 96-bit column sums go to a
//     global scratch plane per word (low, high, overflow count).
//   pass 2 (k_bigcarry): one wave per instance walks the columns in 256-digit
//     tiles, forms t_k = l_k + h_{k-1} + c_{k-2} and resolves the carries with
//     the wave carry look-ahead, storing only the requested window.
// triangles (bigmul_low / bigmul_high, k_bigtri): the low M8 columns of a
//   product, or its top M8 columns (both operands read from their top digit:
//   column k' of the reversed product is column Ktot-1-k' of the natural one),
//   one workgroup per instance, then k_bigcarry resolves the carries and
//   stores the window.  Both passes walk the operands with the running-pointer
//   loop of the uniform multiply (conv_walk_planes).
// windowed add/sub (k_win): r = x*B^dx + y*B^dy (or minus), each operand a
//   window [off, off+len) of a per-instance buffer, optional per-instance
//   predicate (unset: r = x) and carry-in.
// Few instances (the server's division rounds): a lone team or wave per
// instance leaves the device idle and every pass takes as long as one lane's
// walk, so the products split their row blocks over up to 8 workgroups an
// instance (plane set per part, folded by k_sumparts), and the carry pass and
// the windowed add/sub run one workgroup of up to 16 waves per instance
// (k_bigcarry_wide, k_win_wide: chunk carries meet in LDS).
#include <algorithm>

#include "bi_bigmul.h"
#include "bi_kernels.h"
#include "bi_mulcore.h"

namespace bi {

// Rectangle pass 1.  Operands rounded up to 8 digits: A[na8+12] with a_i at
// A[i]; W = A + na8 + 12 with 20 zero words, b_j at W[20+j], zeros up to b
// index nb8+16.
// parts > 1: the row blocks are split over `parts` workgroups per chunk, part
// p adding its rows' products into plane set p (3*K8 words at p*3*K8 of the
// instance's column sums); k_bigcarry adds the sets.
template <int R>
__global__ void __launch_bounds__(1024) k_bigconv(Win a, Win b, uint32_t na8, uint32_t nb8, uint32_t chunks,
                                                  uint32_t parts, uint32_t* __restrict__ cs, uint64_t cs_stride,
                                                  uint32_t K8, uint64_t n) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const uint64_t inst = blockIdx.x / (chunks * parts);
    const uint32_t rest = blockIdx.x % (chunks * parts);
    const uint32_t chunk = rest / parts, part = rest % parts;
    if (inst >= n) return;
    uint32_t* A = lds;
    uint32_t* W = lds + na8 + 12;
    const uint32_t* ga = a.p + inst * a.stride + a.off;
    const uint32_t* gb = b.p + inst * b.stride + b.off;
    for (uint32_t i = threadIdx.x; i < na8 + 12; i += blockDim.x) A[i] = (i < a.len) ? ga[i] : 0u;
    for (uint32_t i = threadIdx.x; i < kWinPad + nb8 + 16; i += blockDim.x) {
        const int64_t j = (int64_t)i - kWinPad;
        W[i] = (j >= 0 && j < (int64_t)b.len) ? gb[j] : 0u;
    }
    __syncthreads();
    const uint32_t NA = na8 / 4, NB = nb8 / 4;
    const uint32_t job = chunk * blockDim.x + threadIdx.x;
    LaneSched sc = sched_rect_r<R>(job, NA, NB);
    uint32_t* CL = cs + inst * cs_stride + (size_t)part * 3 * K8;
    uint32_t* CH = CL + K8;
    uint32_t* CC = CH + K8;
    if (parts == 1) {
        const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
        conv_walk_planes<R>(A, W + 16 - 2 * R, sc, S, CL, CH, CC);
        return;
    }
    const uint32_t TB = na8 / R;   // row blocks of the row operand
    uint32_t kz[2], nz;
    sc = clip_rows(sc, part * TB / parts, (part + 1) * TB / parts, kz, nz);
    const uint64_t zero[4] = {0, 0, 0, 0};
    const uint32_t zc[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < nz; i++) store_cols(CL, CH, CC, kz[i], zero, zc);
    const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
    conv_walk_planes<R, true>(A, W + 16 - 2 * R, sc, S, CL, CH, CC);
}

// Triangle of a product: the low M8 columns of (a mod B^M8)(b mod B^M8)
// (rev = 0), or -- with both operands read from their top digit down
// (rev = 1) -- the top M8 columns of the full product, column k' of the
// reversed product being column na+nb-2-k' of the natural one.  One team of
// M8/8 lanes per instance (sched_low_r) walking the operands in LDS
// (A[M8+12], W with 20 zero words then b); column sums go to global planes of
// M8 columns in reversed-product order.
template <int R>
__global__ void __launch_bounds__(1024) k_bigtri(Win a, Win b, uint32_t M8, uint32_t rev, uint32_t parts,
                                                 uint32_t* __restrict__ cs, uint64_t cs_stride, uint64_t n) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    const uint64_t inst = blockIdx.x / parts;
    const uint32_t part = blockIdx.x % parts;
    if (inst >= n) return;
    uint32_t* A = lds;
    uint32_t* W = lds + M8 + 12;
    const uint32_t* ga = a.p + inst * a.stride + a.off;
    const uint32_t* gb = b.p + inst * b.stride + b.off;
    for (uint32_t i = threadIdx.x; i < M8 + 12; i += blockDim.x) {
        uint32_t v = 0;
        if (i < a.len && i < M8) v = rev ? ga[a.len - 1 - i] : ga[i];
        A[i] = v;
    }
    for (uint32_t i = threadIdx.x; i < kWinPad + M8 + 16; i += blockDim.x) {
        const int64_t j = (int64_t)i - kWinPad;
        uint32_t v = 0;
        if (j >= 0 && j < (int64_t)b.len && j < (int64_t)M8) v = rev ? gb[b.len - 1 - j] : gb[j];
        W[i] = v;
    }
    __syncthreads();
    LaneSched sc = {0, 0, 0, 0, 0, 0};
    if (threadIdx.x < M8 / 8) sc = sched_low_r<R>(threadIdx.x, M8 / 4);
    uint32_t* CL = cs + inst * cs_stride + (size_t)part * 3 * M8;
    uint32_t* CH = CL + M8;
    uint32_t* CC = CH + M8;
    if (parts == 1) {
        const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
        conv_walk_planes<R>(A, W + 16 - 2 * R, sc, S, CL, CH, CC);
        return;
    }
    // rows split over `parts` workgroups, part p into plane set p (as k_bigconv)
    const uint32_t TB = M8 / R;
    uint32_t kz[2], nz;
    sc = clip_rows(sc, part * TB / parts, (part + 1) * TB / parts, kz, nz);
    const uint64_t zero[4] = {0, 0, 0, 0};
    const uint32_t zc[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < nz; i++) store_cols(CL, CH, CC, kz[i], zero, zc);
    const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
    conv_walk_planes<R, true>(A, W + 16 - 2 * R, sc, S, CL, CH, CC);
}

// Row-split products: plane set 0 of each instance += sets 1 .. parts-1 (3*Kp
// words each, columns 4 at a time, Kp % 4 == 0).  A column's (low, high,
// count) triples are summed in 64 bits and folded back to one triple of the
// same value.  One thread per 4 columns: every column at once, where the carry
// pass walks an instance's columns tile after tile.
__global__ void __launch_bounds__(256) k_sumparts(uint32_t* __restrict__ cs, uint64_t cs_stride, uint32_t Kp,
                                                  uint32_t parts, uint64_t n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t per = Kp / 4;
    const uint64_t inst = t / per;
    if (inst >= n) return;
    const uint32_t q = 4 * (uint32_t)(t % per);
    uint32_t* CL = cs + inst * cs_stride;
    uint64_t L[4] = {0, 0, 0, 0}, H[4] = {0, 0, 0, 0};
    uint32_t C[4] = {0, 0, 0, 0};
    for (uint32_t p = 0; p < parts; p++) {
        const uint32_t* P = CL + (size_t)p * 3 * Kp;
        const u32x4 l = *reinterpret_cast<const u32x4*>(P + q);
        const u32x4 h = *reinterpret_cast<const u32x4*>(P + Kp + q);
        const u32x4 c = *reinterpret_cast<const u32x4*>(P + 2 * Kp + q);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            L[j] += l[j];
            H[j] += h[j];
            C[j] += c[j];
        }
    }
    u32x4 l4, h4, c4;
#pragma unroll
    for (int j = 0; j < 4; j++) {   // L + H 2^32 + C 2^64 as one triple
        const uint64_t hh = H[j] + (L[j] >> 32);
        l4[j] = (uint32_t)L[j];
        h4[j] = (uint32_t)hh;
        c4[j] = C[j] + (uint32_t)(hh >> 32);
    }
    *reinterpret_cast<u32x4*>(CL + q) = l4;
    *reinterpret_cast<u32x4*>(CL + Kp + q) = h4;
    *reinterpret_cast<u32x4*>(CL + 2 * Kp + q) = c4;
}

static hipError_t sum_parts(uint32_t* cs, uint64_t cs_stride, uint32_t Kp, uint32_t parts, uint64_t n,
                            hipStream_t st) {
    if (parts <= 1) return hipSuccess;
    const uint64_t threads = n * (Kp / 4);
    hipLaunchKernelGGL(k_sumparts, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, st, cs, cs_stride, Kp, parts,
                       n);
    return hipGetLastError();
}

// Digits [lo, lo+len) of the column sums' value, the carry chain starting at
// natural column k0 (columns below k0 are taken as zero).  Natural column k
// lives in plane column q = k - base (forward) or base - k (reversed);
// columns outside [0, Kp) are zero.  One wave per instance, 256-digit tiles
// starting at k0; a lane reads its 4 columns of each plane as one 16-byte
// load (k0 is chosen so that plane groups are 4-aligned) and takes the high
// words / counts of the columns just below from the lane before it.
// neg (lo == 0 only): store (B^len - value) mod B^len instead, digit k being
// ~v_k + [v_j == 0 for all j < k] (a prefix AND over the wave's ballots).
__device__ __forceinline__ void load_cols(const uint32_t* CL, uint32_t Kp, int64_t base, uint32_t rev, uint32_t k,
                                          u32x4& l4, u32x4& h4, u32x4& c4) {
    const uint32_t* CH = CL + Kp;
    const uint32_t* CC = CH + Kp;
    l4 = h4 = c4 = u32x4{0u, 0u, 0u, 0u};
    if (!rev) {
        const int64_t q = (int64_t)k - base;
        if (q >= 0 && q + 3 < (int64_t)Kp) {
            l4 = *reinterpret_cast<const u32x4*>(CL + q);
            h4 = *reinterpret_cast<const u32x4*>(CH + q);
            c4 = *reinterpret_cast<const u32x4*>(CC + q);
        }
    } else {
        const int64_t q = base - (int64_t)k - 3;   // plane columns q .. q+3 = natural k+3 .. k
        if (q >= 0 && q + 3 < (int64_t)Kp) {
            const u32x4 a = *reinterpret_cast<const u32x4*>(CL + q);
            const u32x4 b = *reinterpret_cast<const u32x4*>(CH + q);
            const u32x4 c = *reinterpret_cast<const u32x4*>(CC + q);
            l4 = u32x4{a[3], a[2], a[1], a[0]};
            h4 = u32x4{b[3], b[2], b[1], b[0]};
            c4 = u32x4{c[3], c[2], c[1], c[0]};
        }
    }
}

__global__ void __launch_bounds__(256) k_bigcarry(const uint32_t* __restrict__ cs, uint64_t cs_stride, uint32_t Kp,
                                                  int64_t base, uint32_t rev, uint32_t k0,
                                                  uint32_t* __restrict__ out, uint64_t out_stride, uint32_t lo,
                                                  uint32_t len, uint64_t n, uint32_t neg) {
    const uint32_t lane = lane_id();
    const uint64_t inst = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (inst >= n) return;
    const uint32_t* CL = cs + inst * cs_stride;
    uint32_t* o = out + inst * out_stride;
    const uint32_t end = lo + len;
    uint32_t carry = 0;
    bool allz = true;   // every digit below the tile is zero (neg)
    // last lane's top values of the previous tile: h_{k-1}, c_{k-2}, c_{k-1}, hi part of t_{k-1}
    uint32_t ph3 = 0, pc2 = 0, pc3 = 0, phi = 0;
    for (uint32_t q0 = k0; q0 < end; q0 += 256) {
        const uint32_t k = q0 + 4 * lane;   // natural columns k .. k+3
        u32x4 l4, h4, c4;
        load_cols(CL, Kp, base, rev, k, l4, h4, c4);
        // the lane below: its h of column k-1 and c of columns k-2, k-1
        uint32_t bh3 = __shfl_up(h4[3], 1, kWave), bc2 = __shfl_up(c4[2], 1, kWave), bc3 = __shfl_up(c4[3], 1, kWave);
        if (lane == 0) {
            bh3 = ph3;
            bc2 = pc2;
            bc3 = pc3;
        }
        uint32_t lo_[4], hi_[4];
        tsplit(l4[0], bh3, bc2, lo_[0], hi_[0]);
        tsplit(l4[1], h4[0], bc3, lo_[1], hi_[1]);
        tsplit(l4[2], h4[1], c4[0], lo_[2], hi_[2]);
        tsplit(l4[3], h4[2], c4[1], lo_[3], hi_[3]);
        uint32_t bhi = __shfl_up(hi_[3], 1, kWave);
        if (lane == 0) bhi = phi;
        const u32x4 X = {lo_[0], lo_[1], lo_[2], lo_[3]};
        const u32x4 Y = {bhi, hi_[0], hi_[1], hi_[2]};
        u32x4 sv;
        bool g, pr;
        rec_lane(X, Y, 0u, sv, g, pr);
        uint32_t cout;
        const uint32_t e = wave_carry(g, pr, carry, cout);
        carry = cout;
        u32x4 r = rec_post(sv, 0u, e);
        if (neg) {
            const uint64_t nzb = ballot((r[0] | r[1] | r[2] | r[3]) != 0u);
            bool z = allz && (nzb & ((1ull << lane) - 1)) == 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t v = r[j];
                r[j] = ~v + (z ? 1u : 0u);
                z = z && v == 0u;
            }
            allz = allz && nzb == 0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t kk = k + j;
            if (kk >= lo && kk < end) o[kk - lo] = r[j];
        }
        ph3 = __shfl(h4[3], kWave - 1, kWave);
        pc2 = __shfl(c4[2], kWave - 1, kWave);
        pc3 = __shfl(c4[3], kWave - 1, kWave);
        phi = __shfl(hi_[3], kWave - 1, kWave);
    }
}

// k_bigcarry for few instances: one workgroup per instance, wave w owning
// tiles [w*tpw, (w+1)*tpw) from k0 (tpw <= kCarryTiles).  Every tile is
// independent but for the carry bit: lane 0 reads the three columns below the
// tile itself (what the tile before would have passed on).  Each wave runs
// its tiles for carry-in 0 and 1, leaves both carry-outs in LDS, resolves its
// carry-in after a barrier and finishes its tiles from registers; neg adds a
// second round for "every digit below is zero".
constexpr int kCarryTiles = 8;

__global__ void __launch_bounds__(1024) k_bigcarry_wide(const uint32_t* __restrict__ cs, uint64_t cs_stride,
                                                        uint32_t Kp, int64_t base, uint32_t rev, uint32_t k0,
                                                        uint32_t* __restrict__ out, uint64_t out_stride, uint32_t lo,
                                                        uint32_t len, uint64_t n, uint32_t neg, uint32_t tpw) {
    __shared__ uint32_t cw[16], nzw[16];
    const uint32_t lane = lane_id(), wv = threadIdx.x / kWave;
    const uint64_t inst = blockIdx.x;
    if (inst >= n) return;   // uniform over the workgroup
    const uint32_t* CL = cs + inst * cs_stride;
    uint32_t* o = out + inst * out_stride;
    const uint32_t end = lo + len;
    u32x4 sv[kCarryTiles];
    bool gg[kCarryTiles], pp[kCarryTiles];
    uint32_t c0 = 0, c1 = 1;
#pragma unroll
    for (int t = 0; t < kCarryTiles; t++) {
        sv[t] = u32x4{0u, 0u, 0u, 0u};
        gg[t] = pp[t] = false;
        const uint32_t q0 = k0 + (wv * tpw + t) * 256;
        if ((uint32_t)t >= tpw || q0 >= end) continue;
        const uint32_t k = q0 + 4 * lane;
        u32x4 l4, h4, c4, pl = {0u, 0u, 0u, 0u}, ph = {0u, 0u, 0u, 0u}, pc = {0u, 0u, 0u, 0u};
        load_cols(CL, Kp, base, rev, k, l4, h4, c4);
        if (lane == 0 && q0 > k0) load_cols(CL, Kp, base, rev, q0 - 4, pl, ph, pc);   // columns q0-4 .. q0-1
        uint32_t bh3 = __shfl_up(h4[3], 1, kWave), bc2 = __shfl_up(c4[2], 1, kWave), bc3 = __shfl_up(c4[3], 1, kWave);
        if (lane == 0) {
            bh3 = ph[3];
            bc2 = pc[2];
            bc3 = pc[3];
        }
        uint32_t lo_[4], hi_[4];
        tsplit(l4[0], bh3, bc2, lo_[0], hi_[0]);
        tsplit(l4[1], h4[0], bc3, lo_[1], hi_[1]);
        tsplit(l4[2], h4[1], c4[0], lo_[2], hi_[2]);
        tsplit(l4[3], h4[2], c4[1], lo_[3], hi_[3]);
        uint32_t bhi = __shfl_up(hi_[3], 1, kWave);
        if (lane == 0) {
            uint32_t plo;
            tsplit(pl[3], ph[2], pc[1], plo, bhi);
        }
        const u32x4 X = {lo_[0], lo_[1], lo_[2], lo_[3]};
        const u32x4 Y = {bhi, hi_[0], hi_[1], hi_[2]};
        rec_lane(X, Y, 0u, sv[t], gg[t], pp[t]);
        uint32_t o0, o1;
        (void)wave_carry(gg[t], pp[t], c0, o0);
        (void)wave_carry(gg[t], pp[t], c1, o1);
        c0 = o0;
        c1 = o1;
    }
    if (lane == 0) cw[wv] = c0 | (c1 << 1);
    __syncthreads();
    uint32_t carry = 0;
    for (uint32_t w = 0; w < wv; w++) carry = (cw[w] >> carry) & 1u;
    uint64_t nz = 0;
#pragma unroll
    for (int t = 0; t < kCarryTiles; t++) {
        const uint32_t q0 = k0 + (wv * tpw + t) * 256;
        if ((uint32_t)t >= tpw || q0 >= end) continue;
        uint32_t cout;
        const uint32_t e = wave_carry(gg[t], pp[t], carry, cout);
        carry = cout;
        sv[t] = rec_post(sv[t], 0u, e);
        nz |= ballot((sv[t][0] | sv[t][1] | sv[t][2] | sv[t][3]) != 0u);
    }
    bool allz = true;   // every digit below the tile is zero (neg)
    if (neg) {
        if (lane == 0) nzw[wv] = nz != 0;
        __syncthreads();
        for (uint32_t w = 0; w < wv; w++) allz = allz && nzw[w] == 0;
    }
#pragma unroll
    for (int t = 0; t < kCarryTiles; t++) {
        const uint32_t q0 = k0 + (wv * tpw + t) * 256;
        if ((uint32_t)t >= tpw || q0 >= end) continue;
        const uint32_t k = q0 + 4 * lane;
        u32x4 r = sv[t];
        if (neg) {
            const uint64_t nzb = ballot((r[0] | r[1] | r[2] | r[3]) != 0u);
            bool z = allz && (nzb & ((1ull << lane) - 1)) == 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t v = r[j];
                r[j] = ~v + (z ? 1u : 0u);
                z = z && v == 0u;
            }
            allz = allz && nzb == 0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t kk = k + j;
            if (kk >= lo && kk < end) o[kk - lo] = r[j];
        }
    }
}

// The carry pass of a long product: one wave per instance, or for few
// instances with many tiles one workgroup of up to 16 waves per instance.
static hipError_t carry_pass(const uint32_t* cs, uint64_t cs_stride, uint32_t Kp, int64_t base, uint32_t rev,
                             uint32_t k0, uint32_t* out, uint64_t out_stride, uint32_t lo, uint32_t len, uint64_t n,
                             uint32_t neg, hipStream_t st) {
    const uint32_t end = lo + len;
    const uint32_t tiles = end > k0 ? (end - k0 + 255) / 256 : 0;
    if (n <= 256 && tiles > 4 && tiles <= 16 * kCarryTiles) {
        const uint32_t nw = std::min<uint32_t>(16, tiles);
        const uint32_t tpw = (tiles + nw - 1) / nw;
        hipLaunchKernelGGL(k_bigcarry_wide, dim3((uint32_t)n), dim3(nw * kWave), 0, st, cs, cs_stride, Kp, base, rev,
                           k0, out, out_stride, lo, len, n, neg, tpw);
    } else {
        hipLaunchKernelGGL(k_bigcarry, dim3((uint32_t)((n + 3) / 4)), dim3(256), 0, st, cs, cs_stride, Kp, base, rev,
                           k0, out, out_stride, lo, len, n, neg);
    }
    return hipGetLastError();
}

void bigmul_prepare() {
    // once per process, thread-safe (static initialisation)
    static const bool attr = [] {
        set_max_lds(reinterpret_cast<const void*>(&k_bigconv<4>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_bigconv<8>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_bigtri<4>), 160 * 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_bigtri<8>), 160 * 1024);
        return true;
    }();
    (void)attr;
}

// Row-split parts of one product at small counts: enough workgroups for one
// per CU (a lone team per instance leaves the device nearly idle, and a
// product then takes as long as one lane's walk), at most 8, each part at
// least 16 row blocks (shorter walks gain less than the extra summing launch
// and the operands' staging in every part cost), within the column-sum
// scratch (words_per_part each).  Measured, 16 divisions of 4096/2048 digits:
// 0.75 -> 0.66 ms; 1024/512 and 256/128 digits, 64 or 256 divisions: unchanged.
static uint32_t row_parts(uint64_t n, uint64_t per_inst, uint32_t row_blocks, uint64_t words_per_part,
                          const BigEnv& env) {
    const uint64_t have = n * per_inst, want = env.num_cus;
    if (have == 0 || have >= want) return 1;
    uint64_t p = (want + have - 1) / have;
    p = std::min<uint64_t>(p, 8);
    p = std::min<uint64_t>(p, row_blocks / 16);
    p = std::min<uint64_t>(p, env.cs_words / (words_per_part * n));
    return p > 1 ? (uint32_t)p : 1u;
}

// Low digits [0, K) of a*b: the triangle of columns < roundup8(K).
hipError_t bigmul_low(Win a, Win b, uint32_t* out, uint64_t out_stride, uint32_t K, uint64_t n, const BigEnv& env,
                      bool neg) {
    if (n == 0 || K == 0) return hipSuccess;
    const uint32_t M8 = (K + 7) & ~7u;
    if (M8 / 8 > 1024) return bigmul(a, b, out, out_stride, 0, K, n, env, neg);   // team too large
    if (env.cs_words < 3ull * M8 * n) return hipErrorInvalidValue;
    const uint32_t parts = row_parts(n, 1, M8 / (uint32_t)env.rows, 3ull * M8, env);
    const uint64_t cs_stride = 3ull * M8 * parts;
    const size_t lds = (size_t)(2 * M8 + 12 + kWinPad + 16) * 4;
    bigmul_prepare();
    const uint32_t block = ((M8 / 8) + 63) / 64 * 64;
    const dim3 grid((uint32_t)(n * parts));
    if (env.rows == 4)
        hipLaunchKernelGGL(k_bigtri<4>, grid, dim3(block), lds, env.st, a, b, M8, 0u, parts, env.cs, cs_stride, n);
    else
        hipLaunchKernelGGL(k_bigtri<8>, grid, dim3(block), lds, env.st, a, b, M8, 0u, parts, env.cs, cs_stride, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = sum_parts(env.cs, cs_stride, M8, parts, n, env.st);
    if (e != hipSuccess) return e;
    return carry_pass(env.cs, cs_stride, M8, (int64_t)0, 0u, 0u, out, out_stride, 0u, K, n, neg ? 1u : 0u, env.st);
}

// Digits [lo, lo+len) of a*b computed from the top columns only: columns
// below lo - 2 are dropped, so the result is the exact window minus 0 or 1
// (the dropped part is below B^lo for any operand length < B).
hipError_t bigmul_high(Win a, Win b, uint32_t* out, uint64_t out_stride, uint32_t lo, uint32_t len, uint64_t n,
                       const BigEnv& env) {
    if (n == 0 || len == 0) return hipSuccess;
    const uint32_t Ktot = a.len + b.len - 1;          // columns 0 .. Ktot-1
    const uint32_t k0 = lo >= 2 ? lo - 2 : 0;
    if (k0 >= Ktot) return bigmul(a, b, out, out_stride, lo, len, n, env);   // zeros (and the top carry)
    const uint32_t M8 = ((Ktot - k0) + 7) & ~7u;
    if (M8 >= Ktot || M8 / 8 > 1024) return bigmul(a, b, out, out_stride, lo, len, n, env);
    if (env.cs_words < 3ull * M8 * n) return hipErrorInvalidValue;
    const uint32_t parts = row_parts(n, 1, M8 / (uint32_t)env.rows, 3ull * M8, env);
    const uint64_t cs_stride = 3ull * M8 * parts;
    const size_t lds = (size_t)(2 * M8 + 12 + kWinPad + 16) * 4;
    bigmul_prepare();
    const uint32_t block = ((M8 / 8) + 63) / 64 * 64;
    const dim3 grid((uint32_t)(n * parts));
    if (env.rows == 4)
        hipLaunchKernelGGL(k_bigtri<4>, grid, dim3(block), lds, env.st, a, b, M8, 1u, parts, env.cs, cs_stride, n);
    else
        hipLaunchKernelGGL(k_bigtri<8>, grid, dim3(block), lds, env.st, a, b, M8, 1u, parts, env.cs, cs_stride, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // reversed column k' holds natural column Ktot-1-k'; start the chain at
    // the lowest natural column the triangle covers
    const uint32_t kstart = Ktot - M8;
    if ((e = sum_parts(env.cs, cs_stride, M8, parts, n, env.st)) != hipSuccess) return e;
    return carry_pass(env.cs, cs_stride, M8, (int64_t)Ktot - 1, 1u, kstart, out, out_stride, lo, len, n, 0u, env.st);
}

hipError_t bigmul(Win a, Win b, uint32_t* out, uint64_t out_stride, uint32_t lo, uint32_t len, uint64_t n,
                  const BigEnv& env, bool neg) {
    if (n == 0 || len == 0) return hipSuccess;
    if (neg && lo != 0) return hipErrorInvalidValue;
    if (a.len < b.len) {
        Win t = a;
        a = b;
        b = t;
    }
    const uint32_t na8 = (a.len + 7) & ~7u, nb8 = ((b.len ? b.len : 1) + 7) & ~7u;
    const uint32_t K8 = na8 + nb8;
    if (env.cs_words < 3ull * K8 * n) return hipErrorInvalidValue;
    const uint32_t NA = na8 / 4;
    // jobs split evenly over the fewest workgroups of <= 1024 lanes
    const uint32_t chunks = (NA + 1023) / 1024;
    const uint32_t block = (((NA + chunks - 1) / chunks) + 63) / 64 * 64;
    const size_t lds = (size_t)(na8 + 12 + kWinPad + nb8 + 16) * 4;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const uint32_t parts = row_parts(n, chunks, na8 / (uint32_t)env.rows, 3ull * K8, env);
    const uint64_t cs_stride = 3ull * K8 * parts;
    bigmul_prepare();
    const dim3 grid((uint32_t)(n * chunks * parts)), blk(block);
    if (env.rows == 4)
        hipLaunchKernelGGL(k_bigconv<4>, grid, blk, lds, env.st, a, b, na8, nb8, chunks, parts, env.cs, cs_stride, K8,
                           n);
    else
        hipLaunchKernelGGL(k_bigconv<8>, grid, blk, lds, env.st, a, b, na8, nb8, chunks, parts, env.cs, cs_stride, K8,
                           n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = sum_parts(env.cs, cs_stride, K8, parts, n, env.st);
    if (e != hipSuccess) return e;
    return carry_pass(env.cs, cs_stride, K8, (int64_t)0, 0u, 0u, out, out_stride, lo, len, n, neg ? 1u : 0u, env.st);
}

size_t bigmul_scratch_words(uint32_t alen, uint32_t blen, uint64_t n) {
    const uint32_t K8 = ((alen + 7) & ~7u) + (((blen ? blen : 1) + 7) & ~7u);
    return 3ull * K8 * n;
}

// ---------------------------------------------------------------------------
// r[k] = X(k - dx) +/- Y(k - dy) (+ cin) for k in [0, len) of each instance,
// X(j) = x[j] for j in [0, x.len) and 0 elsewhere (same for Y).  When pred is
// given and pred[i] == 0, instance i gets r = X (no Y, no carry-in).
template <bool SUB>
__global__ void __launch_bounds__(256) k_win(Win x, int32_t dx, Win y, int32_t dy, uint32_t* __restrict__ r,
                                             uint64_t r_stride, uint32_t len, uint32_t cin,
                                             const uint8_t* __restrict__ pred, uint32_t keep, uint64_t n) {
    const uint32_t lane = lane_id();
    const uint64_t inst = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (inst >= n) return;
    const bool on = pred ? pred[inst] != 0 : true;
    const uint32_t* px = x.p + inst * x.stride + x.off;
    const uint32_t* py = y.p ? y.p + inst * y.stride + y.off : nullptr;
    uint32_t* o = r + inst * r_stride;
    // keep: r is x itself (in place, no shift), so an instance switched off
    // has nothing to write (the delta fix's second pass, and its first for
    // most instances)
    if (!on && keep) return;
    uint32_t carry = on ? cin : 0u;
    for (uint32_t q0 = 0; q0 < len; q0 += 256) {
        uint32_t s4[4], c = 0, all = SUB ? 0u : 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t k = (int64_t)q0 + 4 * lane + j;
            uint32_t xv = 0, yv = 0;
            if (k < len) {
                const int64_t ix = k - dx, iy = k - dy;
                if (ix >= 0 && ix < (int64_t)x.len) xv = px[ix];
                if (on && py && iy >= 0 && iy < (int64_t)y.len) yv = py[iy];
            }
            if (!SUB) {
                const uint64_t t = (uint64_t)xv + yv + c;
                s4[j] = (uint32_t)t;
                c = (uint32_t)(t >> 32);
                all &= s4[j];
            } else {
                const uint64_t t = (uint64_t)xv - yv - c;
                s4[j] = (uint32_t)t;
                c = (uint32_t)(t >> 63);
                all |= s4[j];
            }
        }
        uint32_t cout;
        uint32_t e = wave_carry(c != 0, SUB ? all == 0u : all == 0xFFFFFFFFu, carry, cout);
        carry = cout;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = q0 + 4 * lane + j;
            if (k < len) o[k] = SUB ? s4[j] - e : s4[j] + e;
            e &= SUB ? (uint32_t)(s4[j] == 0u) : (uint32_t)(s4[j] == 0xFFFFFFFFu);
        }
    }
}

// k_win for few long instances: one workgroup per instance, wave w owning
// tiles [w*tpw, (w+1)*tpw) (tpw <= kWinTiles).  Each wave runs its tiles'
// carry chain for carry-in 0 and for carry-in 1 and leaves both carry-outs in
// LDS; after one barrier it resolves its own carry-in from the waves below it
// and replays its tiles (the lane sums and flags are still in registers).  An
// instance then takes about tpw tile steps instead of len/256.  r may alias x
// or y only at the same digit positions (each lane reads its digits before it
// writes them).
constexpr int kWinTiles = 8;

template <bool SUB>
__global__ void __launch_bounds__(1024) k_win_wide(Win x, int32_t dx, Win y, int32_t dy, uint32_t* __restrict__ r,
                                                   uint64_t r_stride, uint32_t len, uint32_t cin,
                                                   const uint8_t* __restrict__ pred, uint32_t keep, uint64_t n,
                                                   uint32_t tpw) {
    __shared__ uint32_t cout_w[16];   // bit 0: carry out for carry-in 0, bit 1: for carry-in 1
    const uint32_t lane = lane_id(), wv = threadIdx.x / kWave;
    const uint64_t inst = blockIdx.x;
    if (inst >= n) return;   // uniform over the workgroup
    const bool on = pred ? pred[inst] != 0 : true;
    const uint32_t* px = x.p + inst * x.stride + x.off;
    const uint32_t* py = y.p ? y.p + inst * y.stride + y.off : nullptr;
    uint32_t* o = r + inst * r_stride;
    if (!on && keep) return;   // in place and switched off: x stays (uniform over the workgroup)
    uint32_t s4[kWinTiles][4];
    bool gl[kWinTiles], pl[kWinTiles];
    uint32_t c0 = 0, c1 = 1;
#pragma unroll
    for (int t = 0; t < kWinTiles; t++) {
        gl[t] = pl[t] = false;
        if ((uint32_t)t >= tpw) continue;
        const uint32_t q0 = (wv * tpw + t) * 256;
        uint32_t c = 0, all = SUB ? 0u : 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t k = (int64_t)q0 + 4 * lane + j;
            uint32_t xv = 0, yv = 0;
            if (k < len) {
                const int64_t ix = k - dx, iy = k - dy;
                if (ix >= 0 && ix < (int64_t)x.len) xv = px[ix];
                if (on && py && iy >= 0 && iy < (int64_t)y.len) yv = py[iy];
            }
            if (!SUB) {
                const uint64_t v = (uint64_t)xv + yv + c;
                s4[t][j] = (uint32_t)v;
                c = (uint32_t)(v >> 32);
                all &= s4[t][j];
            } else {
                const uint64_t v = (uint64_t)xv - yv - c;
                s4[t][j] = (uint32_t)v;
                c = (uint32_t)(v >> 63);
                all |= s4[t][j];
            }
        }
        gl[t] = c != 0;
        pl[t] = SUB ? all == 0u : all == 0xFFFFFFFFu;
        uint32_t o0, o1;
        (void)wave_carry(gl[t], pl[t], c0, o0);
        (void)wave_carry(gl[t], pl[t], c1, o1);
        c0 = o0;
        c1 = o1;
    }
    if (lane == 0) cout_w[wv] = c0 | (c1 << 1);
    __syncthreads();
    uint32_t carry = on ? cin : 0u;
    for (uint32_t w = 0; w < wv; w++) carry = (cout_w[w] >> carry) & 1u;
#pragma unroll
    for (int t = 0; t < kWinTiles; t++) {
        if ((uint32_t)t >= tpw) continue;
        const uint32_t q0 = (wv * tpw + t) * 256;
        uint32_t cout;
        uint32_t e = wave_carry(gl[t], pl[t], carry, cout);
        carry = cout;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = q0 + 4 * lane + j;
            if (k < len) o[k] = SUB ? s4[t][j] - e : s4[t][j] + e;
            e &= SUB ? (uint32_t)(s4[t][j] == 0u) : (uint32_t)(s4[t][j] == 0xFFFFFFFFu);
        }
    }
}

hipError_t win_addsub(bool sub, Win x, int32_t dx, Win y, int32_t dy, uint32_t* r, uint64_t r_stride, uint32_t len,
                      uint32_t cin, const uint8_t* pred, uint64_t n, hipStream_t st) {
    if (n == 0 || len == 0) return hipSuccess;
    // r is x itself: an instance the predicate switches off keeps its digits
    const uint32_t keep = (pred && r == x.p + x.off && r_stride == x.stride && dx == 0 && len <= x.len) ? 1u : 0u;
    const uint32_t tiles = (len + 255) / 256;
    if (n <= 256 && tiles > 4 && tiles <= 16 * kWinTiles) {
        // few long instances: a workgroup of up to 16 waves each
        const uint32_t nw = std::min<uint32_t>(16, tiles);
        const uint32_t tpw = (tiles + nw - 1) / nw;
        const dim3 grid((uint32_t)n), block(nw * kWave);
        if (sub)
            hipLaunchKernelGGL(k_win_wide<true>, grid, block, 0, st, x, dx, y, dy, r, r_stride, len, cin, pred, keep, n,
                               tpw);
        else
            hipLaunchKernelGGL(k_win_wide<false>, grid, block, 0, st, x, dx, y, dy, r, r_stride, len, cin, pred, keep,
                               n, tpw);
        return hipGetLastError();
    }
    const dim3 grid((uint32_t)((n + 3) / 4)), block(256);
    if (sub) hipLaunchKernelGGL(k_win<true>, grid, block, 0, st, x, dx, y, dy, r, r_stride, len, cin, pred, keep, n);
    else hipLaunchKernelGGL(k_win<false>, grid, block, 0, st, x, dx, y, dy, r, r_stride, len, cin, pred, keep, n);
    return hipGetLastError();
}

}  // namespace bi
