// bi_div.hip -- batched division q = floor(u/v), r = u mod v by Newton
// iteration for the whole-shifted inverse, composed from the batched product,
// windowed add/sub, shift and compare kernels, with the whole iteration
// captured in one hipGraph.
//
// Reference: prototypes/div.c:352-467 (shinv + refine + the delta fix of
// Theorem 1), futhark/div.fut:134-167, thesis/sections/division.tex:140-405.
// The reference iterates per instance with data-dependent lengths (h, k) and
// 4m digits of working precision (g = m/4 guard digits, div.c:334).  To give
// every instance of a batch the same kernel shapes, this pipeline first
// normalises each divisor so that its top digit sits at digit NV-1
// (v' = v B^s, u' = u B^s, s = NV-1-findk(v); the quotient is unchanged and
// r = r' / B^s), then runs ONE precision schedule for the whole batch:
//
//   P  = 1 + max_i (prec(u_i) - findk(v_i))   (quotient digits, + 1)
//   X0 = floor(B^6 / (V3 + 1)), V3 = top three digits of v'  (X0 <= B^(NV+3)/v')
//   p -> p2 = min(2p - 2, P + 2):
//       E  = B^(NV+p) - v' X            (>= 0: X never exceeds B^(NV+p)/v')
//       X' = X B^(p2-p) + floor(X (E / B^c) / B^d)      c = max(0, NV-p-2)
//   W  = X at precision P + 2;  relative error <= 4 B^-(P+1)
//   q  = floor(u' W / B^(NV+P+2)) from the top columns of u'W: in {Q-2, .., Q}
//   r' = u' - q v'  (mod B^(NV+1), the low triangle of q v');
//   if r' >= v': q += 1, r' -= v'  (twice)
// Only the digits each step needs are multiplied: the low NV+2 digits of v'X
// (E < B^(NV+2)), the top columns of X E and of u'W (dropping columns two
// below the kept window loses at most one unit, and only downward), and the
// low NV+1 digits of q v'.
//
// Each Newton step squares the relative error (e' <= e^2 + 3 B^-p2), and
// rounding only ever decreases X, so X stays a lower bound of the inverse.
#include <algorithm>
#include <vector>

#include "../../include/bigint.h"
#include "bi_addcore.h"
#include "bi_bigmul.h"
#include "bi_ctx.h"

namespace bi {

// dst[k] = src[k - dir*s_i] for k < dst_len (zero outside src)
__global__ void k_shift_into(const uint32_t* __restrict__ src, uint64_t src_stride, uint32_t src_len,
                             uint32_t* __restrict__ dst, uint64_t dst_stride, uint32_t dst_len,
                             const int32_t* __restrict__ sarr, int dir, uint64_t n) {
    const uint64_t per = ((uint64_t)dst_len + 3) / 4;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= per * n) return;
    const uint64_t i = g / per;
    const uint32_t k0 = (uint32_t)(g - i * per) * 4;
    const int64_t sh = (int64_t)dir * sarr[i];
    const uint32_t* s = src + i * src_stride;
    uint32_t* d = dst + i * dst_stride;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int64_t k = (int64_t)k0 + j;
        if (k < dst_len) {
            const int64_t from = k - sh;
            d[k] = (from >= 0 && from < (int64_t)src_len) ? s[from] : 0u;
        }
    }
}

// per instance: shift to normalise the divisor, zero-divisor flag, and the
// batch maximum of prec(u) - findk(v)
__global__ void k_div_prep(const uint32_t* __restrict__ sigu, const uint32_t* __restrict__ sigv, uint32_t NV,
                           int32_t* __restrict__ s, uint8_t* __restrict__ bad, uint32_t* __restrict__ maxq,
                           uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sv = sigv[i];
    if (sv == 0) {
        s[i] = 0;
        bad[i] = 1;
        return;
    }
    bad[i] = 0;
    const int32_t k = (int32_t)sv - 1;
    s[i] = (int32_t)NV - 1 - k;
    const int32_t qd = (int32_t)sigu[i] - k;
    if (qd > 0) atomicMax(maxq, (uint32_t)qd);
}

// X0 = floor(B^6 / (V3 + 1)) by restoring bit-serial division (4 digits)
__global__ void k_init_recip(const uint32_t* __restrict__ V, uint64_t v_stride, uint32_t NV,
                             uint32_t* __restrict__ X, uint64_t x_stride, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* v = V + i * v_stride;
    const unsigned __int128 D = ((unsigned __int128)v[NV - 1] << 64 | (unsigned __int128)v[NV - 2] << 32 |
                                 (unsigned __int128)v[NV - 3]) + 1;
    // numerator 2^192: its only set bit is the leading one
    unsigned __int128 R = 0;
    unsigned __int128 Q = 0;   // quotient < 2^128
    for (int bit = 192; bit >= 0; bit--) {
        R = (R << 1) | (unsigned __int128)(bit == 192 ? 1 : 0);
        const bool ge = R >= D;
        if (ge) R -= D;
        if (bit < 128) Q |= (unsigned __int128)(ge ? 1 : 0) << bit;
    }
    uint32_t* x = X + i * x_stride;
    x[0] = (uint32_t)Q;
    x[1] = (uint32_t)(Q >> 32);
    x[2] = (uint32_t)(Q >> 64);
    x[3] = (uint32_t)(Q >> 96);
}

// ge[i] = (X >= Y) for windows of different lengths (one wave per instance)
__global__ void __launch_bounds__(256) k_wcmp_ge(Win x, Win y, uint8_t* __restrict__ ge, uint64_t n) {
    const uint32_t lane = lane_id();
    const uint64_t inst = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (inst >= n) return;
    const uint32_t* px = x.p + inst * x.stride + x.off;
    const uint32_t* py = y.p + inst * y.stride + y.off;
    const uint32_t L = x.len > y.len ? x.len : y.len;
    uint32_t found = 0;
    for (int64_t tb = (int64_t)((L - 1) & ~255u); tb >= 0 && !found; tb -= 256) {
        uint32_t best = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t k = (uint32_t)tb + 4 * lane + j;
            if (k < L) {
                const uint32_t a = k < x.len ? px[k] : 0u, b = k < y.len ? py[k] : 0u;
                best = (a != b) ? k + 1 : best;
            }
        }
        found = readfirst(wave_max_u32(best));
    }
    if (lane == 0) {
        bool r = true;
        if (found) {
            const uint32_t k = found - 1;
            const uint32_t a = k < x.len ? px[k] : 0u, b = k < y.len ? py[k] : 0u;
            r = a > b;
        }
        ge[inst] = r ? 1 : 0;
    }
}

// zero-divisor instances: q = r = 0
__global__ void k_div_bad(const uint8_t* __restrict__ bad, uint32_t* __restrict__ q, uint32_t mu,
                          uint32_t* __restrict__ r, uint32_t mv, uint64_t n) {
    const uint64_t i = blockIdx.x;
    if (i >= n || !bad[i]) return;
    for (uint32_t k = threadIdx.x; k < mu; k += blockDim.x) q[i * mu + k] = 0;
    for (uint32_t k = threadIdx.x; k < mv; k += blockDim.x) r[i * mv + k] = 0;
}

__global__ void k_zero32(uint32_t* p, uint64_t count) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// Instantiated graphs of recent shapes / buffers, least recently used first
// out (a serving loop alternates between a few size classes).
constexpr int kDivGraphCache = 16;
struct DivGraph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t key[13] = {};
    uint64_t stamp = 0;
    double macs = 0, tmacs = 0;
    uint32_t P = 0, steps = 0;
};

struct DivState {
    DivGraph g[kDivGraphCache];
    uint64_t clock = 0;
    // owned device buffers (grow-only)
    void* buf = nullptr;
    size_t buf_bytes = 0;
    double last_macs = 0, last_tmacs = 0;   // digit products in all / of them by transform
    uint32_t last_P = 0, last_steps = 0;
    void drop(DivGraph& x) {
        if (x.exec) (void)hipGraphExecDestroy(x.exec);
        if (x.graph) (void)hipGraphDestroy(x.graph);
        x = DivGraph();
    }
    void drop_all() {
        for (auto& x : g) drop(x);
    }
};

static DivState* div_state(bi_ctx* ctx) {
    if (!ctx->extras) ctx->extras = new DivState();
    return static_cast<DivState*>(ctx->extras);
}

void div_state_free(bi_ctx* ctx) {
    DivState* d = static_cast<DivState*>(ctx->extras);
    if (!d) return;
    d->drop_all();
    if (d->buf) (void)hipFree(d->buf);
    delete d;
    ctx->extras = nullptr;
}

struct DivLayout {
    uint32_t NV, NU, P, PP, XC, EC, YC, QC, RC;
    int rows;   // digits per block step of the long products
    size_t off_U, off_V, off_X0, off_X1, off_E, off_Y, off_Yc, off_Q, off_QV, off_R, off_cs, off_VT;
    size_t cs_words, total_words;
    bool vt;    // V's transform kept for reuse (region off_VT, n * 2^13 points)
};

static bool keep_v_transform(uint32_t NV, uint64_t n);

static DivLayout div_layout(uint32_t mu, uint32_t mv, uint32_t P, uint64_t n) {
    DivLayout L;
    L.NV = std::max<uint32_t>(mv, 3);
    L.NU = mu + L.NV;
    L.P = P;
    L.PP = P + 2;
    L.XC = L.PP + 2;
    L.EC = L.NV + L.PP + 1;
    L.YC = L.PP + 1 + L.NV + 3;
    L.QC = P;
    L.RC = L.NV + 1;
    size_t o = 0;
    auto take = [&](size_t words) {
        const size_t r = o;
        o += (words * n + 63) & ~(size_t)63;
        return r;
    };
    L.off_U = take(L.NU);
    L.off_V = take(L.NV);
    L.off_X0 = take(L.XC);
    L.off_X1 = take(L.XC);
    L.off_E = take(L.EC);
    L.off_Y = take(L.YC);
    L.off_Yc = take(L.PP + 2);
    L.off_Q = take(L.QC);
    L.off_QV = take(L.RC);
    L.off_R = take(L.RC);
    // the largest product: u' (P+NV-2 digits) x W (PP+1 digits)
    const uint32_t LU = std::max<uint32_t>(1, std::min(L.NU, P + L.NV - 2));
    size_t cs = bigmul_scratch_words(LU, L.PP + 1, 1);
    cs = std::max(cs, bigmul_scratch_words(L.NV, L.PP + 1, 1));
    cs = std::max(cs, bigmul_scratch_words(L.PP + 1, L.NV + 3, 1));
    cs = std::max(cs, bigmul_scratch_words(L.QC, L.NV, 1));
    // room for the long products' row-split parts at small counts (up to 8
    // plane sets an instance, within what 256 instances take -- the size the
    // server reserves)
    const uint64_t cs_inst = std::max<uint64_t>(n, std::min<uint64_t>(8 * n, 256));
    L.cs_words = cs * cs_inst;
    L.off_cs = o;
    o += (cs * cs_inst + 63) & ~(size_t)63;
    // V's 2^13-point transform (8-byte points) for the products that reuse it
    L.vt = keep_v_transform(L.NV, n);
    L.off_VT = o;
    if (L.vt) o += (size_t)n * 2 * (1u << 13);
    L.total_words = o;
    return L;
}

// The Newton precisions after the initial 3 digits, in order (at most 64):
// a step from p digits may reach at most 2p - 2, so they are chosen from the
// end -- PP, then (q + 3) / 2 (= ceil((q + 2) / 2)) of the previous one, down
// to 3.  That reaches PP in as many steps as doubling from 3 (3, 4, 6, 10,
// ..., 2p - 2) but never ends with a short step: doubling can stop one digit
// short of PP and then spends a whole step -- a V X product of full length --
// on it.
static int newton_schedule(uint32_t PP, uint32_t* sched) {
    uint32_t rev[64];
    int ns = 0;
    for (uint32_t q = PP; q > 3 && ns < 64; q = (q + 3) / 2) rev[ns++] = q;
    for (int i = 0; i < ns; i++) sched[i] = rev[ns - 1 - i];
    return ns;
}

// V X of a Newton step by transform: where the transform product costs
// clearly less than the quadratic rectangle (2^13 points ~ 3.7e6 quadratic
// multiply-adds, 2^14 points ~ 7.5e6; taken from about 1.6 times that)
// words an instance of the wrap-around remainder takes in the column-sum
// scratch: the cyclic product's digits, then k_wrap_rem's working copy
constexpr uint64_t kWrapWords = 2ull * ((kCyclicDigits + 4 + 3) & ~3u);

// Quotients of at least this many digits (+ 1) come in two halves from an
// inverse of about half the precision (div_record).
constexpr uint32_t kSplitMinP = 64;

// The remainder by the wrap-around identity (k_wrap_rem) where the divisor is
// long enough for the transform to beat the low triangle, every part fits
// its 6400-digit window and the batch fills the device (a transform is one
// workgroup per instance; at small counts the row-split triangle is faster:
// 16 divisions of 4096/2048 digits 0.54 ms against 0.63 ms).
static bool rem_by_wrap(uint32_t NV, uint32_t NU, uint32_t P, uint64_t n) {
    return n >= 512 && NV >= 2048 && NV + 1 < kCyclicDigits && P <= kCyclicDigits && NU <= 2 * kCyclicDigits;
}

// V's transform (2^13 points, 25-bit chunks -- the shape of the cyclic
// remainder's product) is kept across the batch's transform products with V
// where the batch is large enough for the transform products at all: each
// product that reuses it skips one of its three transforms, for the cost of
// writing it once and reading it back (n * 64 KiB each way).
static bool keep_v_transform(uint32_t NV, uint64_t n) {
    return n >= 512 && NV >= 2048 && NV + 1 < kCyclicDigits;
}

// A Newton step's V X (NV x p1 digits) by transform rather than quadratic:
// a 2^13-point transform product costs about as much as 4.2e6 quadratic
// digit products (one MI355X, 16384 instances), 2/3 of that when V's
// transform is reused, so from 4e6 where V's transform is kept (the first
// such product computes and stores it) and from 6e6 otherwise; 2^14 points
// (two transforms of twice the length) from 12e6.
static bool vx_by_transform(uint32_t NV, uint32_t p1, uint64_t n) {
    const NttPlan pl = ntt_plan(p1, NV);
    const double rect = (double)NV * p1;
    const bool keep = keep_v_transform(NV, n) && pl.w == 25;
    return (pl.lg == 12 && pl.w == 25 && rect >= 2e6) || (pl.lg == 13 && rect >= (keep ? 4e6 : 6e6)) ||
           (pl.lg == 14 && rect >= 12e6);
}


// The remainder U - Q V by the wrap-around identity.  0 <= U - Q V < B^rlen
// (Q is at most the true quotient and the delta fix leaves less than 3V),
// and with N = kCyclicDigits > rlen it is the residue of
//   U mod B^N + floor(U / B^N) - Q V   mod (B^N - 1),
// Q V mod (B^N - 1) being the value C of the cyclic transform (digits
// [0, N+3), the top three still to be folded back).  One wave per instance,
// one-bit carry look-aheads throughout:
//   1. D = (U_L + U_H) - C over N + 3 digits, two's complement (an add chain,
//      then a subtract chain, per tile); its top h = D div B^N is small;
//   2. E = (D mod B^N) + h, signed, with carry or borrow w (B^N = 1 again);
//   3. F = E + w, which cannot carry again (E is small when w = +1 and
//      large when w = -1);
//   F = B^N - 1 stands for 0.  T (N + 4 words an instance) holds D, E, F.
__global__ void __launch_bounds__(256) k_wrap_rem(const uint32_t* __restrict__ U, uint64_t u_stride, uint32_t NU,
                                                  const uint32_t* __restrict__ C, uint32_t* __restrict__ T,
                                                  uint64_t ct_stride, uint32_t* __restrict__ R, uint64_t r_stride,
                                                  uint32_t rlen, uint64_t n) {
    constexpr uint32_t N = kCyclicDigits;
    static_assert(N % 256 == 0, "digit N starts a tile");
    const uint32_t lane = lane_id();
    const uint64_t inst = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (inst >= n) return;
    const uint32_t* u = U + inst * u_stride;   // NU digits
    const uint32_t* c = C + inst * ct_stride;
    uint32_t* t = T + inst * ct_stride;
    uint32_t* r = R + inst * r_stride;
    // Every pass gives each lane the same digits (4 per tile), so T is only
    // ever read back by the lane that wrote it; values other lanes need
    // travel by readlane.
    // 1. D = U_L + U_H - C over N + 3 digits; h = D div B^N from digits N, N+1
    //    (lane 0 of the tile at N; |h| < 2^42, digit N+2 is its sign)
    uint32_t ca = 0, cb = 0, hlo = 0, hhi = 0;
    for (uint32_t q0 = 0; q0 < N + 3; q0 += 256) {
        const uint32_t k = q0 + 4 * lane;
        u32x4 x, y, z;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t kk = k + j;
            x[j] = (kk < N && kk < NU) ? u[kk] : 0u;
            y[j] = (kk < N && N + kk < NU) ? u[N + kk] : 0u;
            z[j] = kk < N + 3 ? c[kk] : 0u;
        }
        u32x4 s;
        bool g, p;
        lane_pre<false>(x, y, 0u, s, g, p);
        uint32_t cout;
        s = lane_post<false>(s, 0u, wave_carry(g, p, ca, cout));
        ca = cout;
        lane_pre<true>(s, z, 0u, s, g, p);
        s = lane_post<true>(s, 0u, wave_carry(g, p, cb, cout));
        cb = cout;
        if (q0 < N) {
            *reinterpret_cast<u32x4*>(t + k) = s;
        } else {
            hlo = __builtin_amdgcn_readfirstlane(s[0]);
            hhi = __builtin_amdgcn_readfirstlane(s[1]);
        }
    }
    const int64_t h = (int64_t)(((uint64_t)hhi << 32) | hlo);
    const bool hneg = h < 0;
    const uint64_t ha = hneg ? (uint64_t)(-h) : (uint64_t)h;
    // 2. E = D mod B^N + h (carry or borrow w)   3. F = E + w
    int w = 0;
    uint32_t top = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && w == 0) break;
        const bool sub = pass == 0 ? hneg : (w < 0);
        uint32_t cc = 0;
        for (uint32_t q0 = 0; q0 < N; q0 += 256) {
            const uint32_t k = q0 + 4 * lane;
            const u32x4 x = *reinterpret_cast<const u32x4*>(t + k);
            u32x4 y = {0u, 0u, 0u, 0u};
            if (k == 0) {
                y[0] = pass == 0 ? (uint32_t)ha : 1u;
                y[1] = pass == 0 ? (uint32_t)(ha >> 32) : 0u;
            }
            u32x4 s;
            bool g, p;
            uint32_t cout;
            if (sub) {
                lane_pre<true>(x, y, 0u, s, g, p);
                s = lane_post<true>(s, 0u, wave_carry(g, p, cc, cout));
            } else {
                lane_pre<false>(x, y, 0u, s, g, p);
                s = lane_post<false>(s, 0u, wave_carry(g, p, cc, cout));
            }
            cc = cout;
            *reinterpret_cast<u32x4*>(t + k) = s;
            if (q0 + 256 == N) top = __builtin_amdgcn_readlane(s[3], kWave - 1);   // digit N-1
        }
        if (pass == 0) w = cc ? (sub ? -1 : 1) : 0;
    }
    // F = B^N - 1 (top digit all ones; a remainder below B^rlen has it zero)
    // stands for 0; each lane copies its own digits of F
    const bool zero = top == 0xFFFFFFFFu;
    for (uint32_t q0 = 0; q0 < rlen; q0 += 256) {
        const uint32_t k = q0 + 4 * lane;
        const u32x4 f = *reinterpret_cast<const u32x4*>(t + k);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (k + j < rlen) r[k + j] = zero ? 0u : f[j];
    }
}

// The quotient's digits by the same identity.  With U' = U[i0, LU) and P' =
// U' W (ma + mb digits, h = ma + mb - N of them above B^N) and L1 = P' mod
// B^h (a low triangle), P' is fixed by its residues modulo B^N - 1 (the
// cyclic transform's C) and modulo B^h (gcd 1, and P' < (B^N - 1) B^h):
// P' = L1 + B^h Y with Y = (C - L1) B^(-h) mod (B^N - 1), and as B^(-h) is a
// digit rotation, digit k >= h of P' is digit k mod N of Z = (C - L1) mod
// (B^N - 1).  One wave per instance:
//   1. D = C - L1 over N + 3 digits, two's complement; hv = D div B^N;
//   2. E = (D mod B^N) + hv with carry or borrow w;   3. F = E + w;
//   Z = F, or 0 where F = B^N - 1 (every digit all ones);
//   4. out[t] = digit (first + t) mod N of Z for t < len (first >= h), 0 from
//      digit h + N on (P' < B^(h+N)).
// T (N + 4 words an instance) holds D, E, F.
__global__ void __launch_bounds__(256) k_wrap_quot(const uint32_t* __restrict__ C, uint32_t* __restrict__ T,
                                                   uint64_t ct_stride, const uint32_t* __restrict__ L1, uint32_t h,
                                                   uint64_t l_stride, uint32_t* __restrict__ out, uint64_t o_stride,
                                                   uint32_t first, uint32_t len, uint64_t n) {
    constexpr uint32_t N = kCyclicDigits;
    static_assert(N % 256 == 0, "digit N starts a tile");
    const uint32_t lane = lane_id();
    const uint64_t inst = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (inst >= n) return;
    const uint32_t* c = C + inst * ct_stride;
    const uint32_t* l1 = L1 + inst * l_stride;
    uint32_t* t = T + inst * ct_stride;
    uint32_t* o = out + inst * o_stride;
    // 1. D = C - L1; digits N, N+1 (lane 0 of the tile at N) give hv
    //    (|hv| < 2^42, digit N+2 its sign)
    uint32_t cb = 0, hlo = 0, hhi = 0;
    for (uint32_t q0 = 0; q0 < N + 3; q0 += 256) {
        const uint32_t k = q0 + 4 * lane;
        u32x4 x, y;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t kk = k + j;
            x[j] = kk < N + 3 ? c[kk] : 0u;
            y[j] = kk < h ? l1[kk] : 0u;
        }
        u32x4 s;
        bool g, p;
        uint32_t cout;
        lane_pre<true>(x, y, 0u, s, g, p);
        s = lane_post<true>(s, 0u, wave_carry(g, p, cb, cout));
        cb = cout;
        if (q0 < N) {
            *reinterpret_cast<u32x4*>(t + k) = s;
        } else {
            hlo = __builtin_amdgcn_readfirstlane(s[0]);
            hhi = __builtin_amdgcn_readfirstlane(s[1]);
        }
    }
    const int64_t hv = (int64_t)(((uint64_t)hhi << 32) | hlo);
    const bool hneg = hv < 0;
    const uint64_t ha = hneg ? (uint64_t)(-hv) : (uint64_t)hv;
    // 2. E = D mod B^N + hv (carry or borrow w)   3. F = E + w; every digit
    //    of the last pass written is folded into `ones` (F = B^N - 1 test)
    int w = 0;
    uint32_t ones = 0xFFFFFFFFu;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && w == 0) break;
        const bool sub = pass == 0 ? hneg : (w < 0);
        uint32_t cc = 0;
        ones = 0xFFFFFFFFu;
        for (uint32_t q0 = 0; q0 < N; q0 += 256) {
            const uint32_t k = q0 + 4 * lane;
            const u32x4 x = *reinterpret_cast<const u32x4*>(t + k);
            u32x4 y = {0u, 0u, 0u, 0u};
            if (k == 0) {
                y[0] = pass == 0 ? (uint32_t)ha : 1u;
                y[1] = pass == 0 ? (uint32_t)(ha >> 32) : 0u;
            }
            u32x4 s;
            bool g, p;
            uint32_t cout;
            if (sub) {
                lane_pre<true>(x, y, 0u, s, g, p);
                s = lane_post<true>(s, 0u, wave_carry(g, p, cc, cout));
            } else {
                lane_pre<false>(x, y, 0u, s, g, p);
                s = lane_post<false>(s, 0u, wave_carry(g, p, cc, cout));
            }
            cc = cout;
            *reinterpret_cast<u32x4*>(t + k) = s;
            ones &= s[0] & s[1] & s[2] & s[3];
        }
        if (pass == 0) w = cc ? (sub ? -1 : 1) : 0;
    }
    const bool zero = ballot(ones == 0xFFFFFFFFu) == ~0ull;
    // 4. the rotated window of Z (T was written by other lanes: wave fence)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t q0 = 0; q0 < len; q0 += 256) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t i = q0 + 4 * lane + j;
            if (i < len) {
                // digits at and above h + N are past the product (Y < B^N)
                const uint32_t k = first + i;
                const uint32_t kz = k >= N ? k - N : k;   // first + len <= 2N
                o[i] = (zero || k >= h + N) ? 0u : t[kz];
            }
        }
    }
}

// The quotient by k_wrap_quot where the shapes fit the 6400-digit transform
// window and the batch fills the device: U' (numerator digits from NV - 1
// up) and W at most N digits each, the wrapped part h no longer than the
// quotient's offset (the window reads digits above it only), and the
// transform product (about 4.2e6 quadratic digit products' time, see
// vx_by_transform) plus the low triangle (h^2/2) under 0.9 of the top
// triangle it replaces (about ma^2/2).
static bool quot_by_wrap(uint32_t NV, uint32_t LU, uint32_t PP, uint64_t n) {
    if (n < 512 || NV < 2048 || LU <= NV) return false;
    const uint32_t ma = LU - (NV - 1), mb = PP + 1;
    if (ma > kCyclicDigits || mb > kCyclicDigits || ma + mb <= kCyclicDigits) return false;
    const uint32_t h = ma + mb - kCyclicDigits;
    return h <= PP + 1 && 4.2e6 + 0.5 * (double)h * h < 0.9 * 0.5 * (double)ma * ma;
}

// Records the pipeline (everything after the batch maximum is known) on st.
static hipError_t div_record(bi_ctx* ctx, const DivLayout& L, uint32_t* base, const uint32_t* us, const uint32_t* vs,
                             uint32_t mu, uint32_t mv, uint32_t* qs, uint32_t* rs, const int32_t* d_s,
                             const uint8_t* d_bad, uint8_t* d_flag, uint64_t n, hipStream_t st, BigEnv env,
                             double* macs, double* tmacs, uint32_t* steps) {
    uint32_t* U = base + L.off_U;
    uint32_t* V = base + L.off_V;
    uint32_t* Xb[2] = {base + L.off_X0, base + L.off_X1};
    uint32_t* E = base + L.off_E;
    uint32_t* Y = base + L.off_Y;
    uint32_t* Yc = base + L.off_Yc;
    uint32_t* Q = base + L.off_Q;
    uint32_t* QV = base + L.off_QV;
    uint32_t* R = base + L.off_R;
    uint64_t* VT = reinterpret_cast<uint64_t*>(base + L.off_VT);
    bool vt_ready = false;   // VT holds V's transform
    env.cs = base + L.off_cs;
    env.cs_words = L.cs_words;
    env.st = st;
    env.rows = L.rows;
    const uint32_t NV = L.NV, NU = L.NU, P = L.P, PP = L.PP;
    hipError_t e;
    auto blocks = [](uint64_t x, uint32_t b) { return dim3((uint32_t)((x + b - 1) / b)); };
    // macs: digit products of every long product, per instance (a product
    // taken by transform counts its quadratic equivalent); tmacs: those of
    // the products taken by transform
    *macs = 0;
    *tmacs = 0;
    *steps = 0;

    // normalise: V = v B^s (NV digits), U = u B^s (NU digits)
    hipLaunchKernelGGL(k_shift_into, blocks(((NV + 3) / 4) * n, 256), dim3(256), 0, st, vs, (uint64_t)mv, mv, V,
                       (uint64_t)NV, NV, d_s, 1, n);
    hipLaunchKernelGGL(k_shift_into, blocks(((NU + 3) / 4) * n, 256), dim3(256), 0, st, us, (uint64_t)mu, mu, U,
                       (uint64_t)NU, NU, d_s, 1, n);
    hipLaunchKernelGGL(k_zero32, blocks((uint64_t)L.XC * n, 256), dim3(256), 0, st, Xb[0], (uint64_t)L.XC * n);
    hipLaunchKernelGGL(k_init_recip, blocks(n, 256), dim3(256), 0, st, V, (uint64_t)NV, NV, Xb[0], (uint64_t)L.XC, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    // The quotient in two halves (q = q1 B^k + q0, each from an inverse of
    // about half the precision) where it is long: the last Newton step (a V
    // X of NV x P/2 digits and an X E triangle) and the full-length U W give
    // way to two quotient steps of half the length.  q1 = floor(U2 / V), U2
    // = floor(U / B^k), has at most P - k - 1 digits; r1 = U2 - q1 V < V and
    // U3 = r1 B^k + (U mod B^k) < V B^k give q0 = floor(U3 / V) < B^k (P0 =
    // k + 2 keeps all NV + k digits of U3 in the product).  Each half is a
    // full quotient step of its own (top columns of U W, r' from the low part
    // of Q V, the delta fix), so W needs max(P1, P0) + 2 digits.
    const bool split = P >= kSplitMinP;
    const uint32_t k_lo = split ? (P - 1) / 2 : 0;
    const uint32_t P1 = P - k_lo, P0 = k_lo + 2;
    const uint32_t target = split ? std::max(P1, P0) + 2 : PP;

    // Newton steps (precisions from newton_schedule)
    uint32_t sched[64];
    const int ns = newton_schedule(target, sched);
    uint32_t p = 3;
    int cur = 0;
    for (int si = 0; si < ns; si++) {
        const uint32_t p2 = sched[si];   // p2 <= 2p - 2
        uint32_t* X = Xb[cur];
        uint32_t* X2 = Xb[cur ^ 1];
        // E = B^(NV+p) - V X.  E < B^(NV+2), so only the low NV+2 digits of
        // V X are needed: E = (0 - V X) mod B^(NV+2), negated by the
        // product's own carry pass.  The low digits come from the full
        // (thin) rectangle while p is small, from the low triangle once
        // that is cheaper.  The step reads only E / B^c, and V's digits
        // below cw = c - p - 1 move it by less than one unit there (V_l X <
        // B^(cw+p+1) = B^c): the product takes V's top NV - cw digits only,
        // which gives E_hi = E + V_l X, at most one unit of E / B^c high, so
        // the step's X' is at most one unit high, and X' - 1 keeps X a lower
        // bound (the error grows by at most 2 units).
        const uint32_t c = (NV >= p + 2) ? NV - p - 2 : 0;
        const uint32_t cw = c > p + 1 ? c - p - 1 : 0;
        const uint32_t NVh = NV - cw;
        const double rect = (double)NVh * (p + 1), tri = 0.5 * (double)(NVh + 9) * (NVh + 9);
        if (vx_by_transform(NVh, p + 1, n)) {
            // the whole product by transform, its low NV+2-cw digits stored
            // negated; V second, its transform kept or reused where it has
            // the kept shape (the whole of V)
            const NttPlan pl = ntt_plan(p + 1, NVh);
            const bool keep = cw == 0 && L.vt && pl.lg == 13 && pl.w == 25;
            e = ntt_launch(ctx, n, p + 1, NVh, X, L.XC, V + cw, NV, E + cw, L.EC, 0, NV + 2 - cw, true, st, false,
                           keep && !vt_ready ? VT : nullptr, keep && vt_ready ? VT : nullptr);
            vt_ready |= keep;
            *macs += rect;
            *tmacs += rect;
        } else if (rect <= tri) {
            // digits [cw, NV + 2) of E_hi = -(V_h X) B^cw mod B^(NV+2)
            e = bigmul(Win{V, NV, cw, NVh}, Win{X, L.XC, 0, p + 1}, E + cw, L.EC, 0, NV + 2 - cw, n, env, true);
            *macs += rect;
        } else {
            e = bigmul_low(Win{V, NV, cw, NVh}, Win{X, L.XC, 0, p + 1}, E + cw, L.EC, NV + 2 - cw, n, env, true);
            *macs += tri;
        }
        if (e != hipSuccess) return e;
        const uint32_t lenE = NV + 2 - c;
        const uint32_t d = NV + 2 * p - p2 - c;
        // Y = X * (E / B^c); keep digits [d, d + p2 + 2) -- from the top
        // columns only, which can only lower X' (it stays a lower bound)
        e = bigmul_high(Win{X, L.XC, 0, p + 1}, Win{E, L.EC, c, lenE}, Yc, PP + 2, d, p2 + 2, n, env);
        if (e != hipSuccess) return e;
        {
            const double kr = (double)(p + lenE) - d + 9;
            *macs += std::min((double)(p + 1) * lenE, 0.5 * kr * kr);
        }
        // X' = X B^(p2-p) + Yc  (p2+2 digits computed, top digit is zero)
        e = win_addsub(false, Win{X, L.XC, 0, p + 1}, (int32_t)(p2 - p), Win{Yc, PP + 2, 0, p2 + 2}, 0, X2, L.XC,
                       std::min(p2 + 2, L.XC), 0, nullptr, n, st);
        if (e != hipSuccess) return e;
        if (cw) {   // X' - 1 (X' >= B^p2 > 0)
            e = win_addsub(true, Win{X2, L.XC, 0, std::min(p2 + 2, L.XC)}, 0, Win{nullptr, 0, 0, 0}, 0, X2, L.XC,
                           std::min(p2 + 2, L.XC), 1, nullptr, n, st);
            if (e != hipSuccess) return e;
        }
        cur ^= 1;
        p = p2;
        (*steps)++;
    }
    uint32_t* W = Xb[cur];   // target + 1 digits
    (void)Y;
    // Quotient and remainder of the numerator U's row window [uoff, uoff +
    // ulen) by V from W at precision Ph (>= Pq + 2): Pq - 1 digits into Qd
    // (row stride qs), r' into R, delta-fixed.
    auto quot_rem = [&](uint32_t uoff, uint32_t ulen, uint32_t Ph, uint32_t Pq, uint32_t* Qd,
                        uint64_t qs) -> hipError_t {
        hipError_t e;
        // q = floor(U W / B^(NV+Ph)), digits [NV+Ph, NV+Ph+Pq)
        const uint32_t LU = std::max<uint32_t>(1, std::min(ulen, Pq + NV - 2));
        // (top columns only: q may come out one lower, which the delta fix
        // absorbs; the wrap-around product drops the numerator digits below
        // NV-1, whose products stay below one unit of q, with the same effect)
        const double kr = (double)(LU + Ph + 1) - (NV + Ph) + 9;
        if (quot_by_wrap(NV, LU, Ph, n) && env.cs_words >= n * kWrapWords) {
            const uint32_t i0 = NV - 1, ma = LU - i0, mb = Ph + 1, h = ma + mb - kCyclicDigits;
            // L1 = U' W mod B^h (in Yc, free after the Newton steps), then the
            // cyclic product into the column-sum scratch the triangle no longer uses
            e = bigmul_low(Win{U, NU, uoff + i0, ma}, Win{W, L.XC, 0, mb}, Yc, PP + 2, h, n, env);
            if (e != hipSuccess) return e;
            e = ntt_launch(ctx, n, ma, mb, U + uoff + i0, NU, W, L.XC, env.cs, kWrapWords, 0, kCyclicDigits + 3,
                           false, st, true);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_wrap_quot, blocks(n, 4), dim3(256), 0, st, env.cs, env.cs + kWrapWords / 2,
                               kWrapWords, Yc, h, (uint64_t)(PP + 2), Qd, qs, NV + Ph - i0, Pq, n);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            *macs += 0.5 * kr * kr + 0.5 * (double)h * h;
            *tmacs += 0.5 * kr * kr;
        } else {
            e = bigmul_high(Win{U, NU, uoff, LU}, Win{W, L.XC, 0, Ph + 1}, Qd, qs, NV + Ph, Pq, n, env);
            if (e != hipSuccess) return e;
            *macs += std::min((double)LU * (Ph + 1), 0.5 * kr * kr);
        }
        // r' = U - Q V  (mod B^(NV+1)): the low triangle of Q V, or the whole
        // remainder through the cyclic transform (k_wrap_rem); V's transform
        // is stored by the first such product of the batch that needs it
        if (rem_by_wrap(NV, ulen, Pq, n) && env.cs_words >= n * kWrapWords) {
            *tmacs += 0.5 * (double)(NV + 8) * (NV + 8);
            e = ntt_launch(ctx, n, Pq, NV, Qd, qs, V, NV, env.cs, kWrapWords, 0, kCyclicDigits + 3, false, st, true,
                           L.vt && !vt_ready ? VT : nullptr, L.vt && vt_ready ? VT : nullptr);
            if (e != hipSuccess) return e;
            vt_ready |= L.vt;
            hipLaunchKernelGGL(k_wrap_rem, blocks(n, 4), dim3(256), 0, st, U + uoff, (uint64_t)NU, ulen, env.cs,
                               env.cs + kWrapWords / 2, kWrapWords, R, (uint64_t)L.RC, NV + 1, n);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        } else {
            e = bigmul_low(Win{Qd, qs, 0, Pq}, Win{V, NV, 0, NV}, QV, L.RC, NV + 1, n, env);
            if (e != hipSuccess) return e;
            e = win_addsub(true, Win{U, NU, uoff, std::min(NV + 1, ulen)}, 0, Win{QV, L.RC, 0, NV + 1}, 0, R, L.RC,
                           NV + 1, 0, nullptr, n, st);
            if (e != hipSuccess) return e;
        }
        *macs += 0.5 * (double)(NV + 8) * (NV + 8);
        // delta fix, twice (the second pass is a no-op when the bound holds)
        for (int it = 0; it < 2; it++) {
            hipLaunchKernelGGL(k_wcmp_ge, blocks(n, 4), dim3(256), 0, st, Win{R, L.RC, 0, NV + 1},
                               Win{V, NV, 0, NV}, d_flag, n);
            e = win_addsub(false, Win{Qd, qs, 0, Pq}, 0, Win{nullptr, 0, 0, 0}, 0, Qd, qs, Pq, 1, d_flag, n, st);
            if (e != hipSuccess) return e;
            e = win_addsub(true, Win{R, L.RC, 0, NV + 1}, 0, Win{V, NV, 0, NV}, 0, R, L.RC, NV + 1, 0, d_flag, n,
                           st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    if (!split) {
        e = quot_rem(0, NU, PP, P, Q, L.QC);
        if (e != hipSuccess) return e;
    } else {
        // q1 (P1 - 1 digits) into Q from digit k_lo up, r1 into R
        e = quot_rem(k_lo, NU - k_lo, target, P1, Q + k_lo, L.QC);
        if (e != hipSuccess) return e;
        // the second numerator r1 B^k_lo + (U mod B^k_lo), in place: U's
        // digits from k_lo up become r1 and zeros
        e = win_addsub(false, Win{R, L.RC, 0, NV + 1}, 0, Win{nullptr, 0, 0, 0}, 0, U + k_lo, NU, NU - k_lo, 0,
                       nullptr, n, st);
        if (e != hipSuccess) return e;
        // q0 (< B^k_lo) in E (free after the Newton steps), r into R; then its
        // low k_lo digits into Q
        e = quot_rem(0, NV + 1 + k_lo, target, P0, E, L.EC);
        if (e != hipSuccess) return e;
        e = win_addsub(false, Win{E, L.EC, 0, k_lo}, 0, Win{nullptr, 0, 0, 0}, 0, Q, L.QC, k_lo, 0, nullptr, n, st);
        if (e != hipSuccess) return e;
    }
    // outputs: q (mu digits), r = r' / B^s (mv digits)
    e = win_addsub(false, Win{Q, L.QC, 0, std::min(P, mu)}, 0, Win{nullptr, 0, 0, 0}, 0, qs, mu, mu, 0, nullptr, n,
                   st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_shift_into, blocks(((mv + 3) / 4) * n, 256), dim3(256), 0, st, R, (uint64_t)L.RC, NV + 1, rs,
                       (uint64_t)mv, mv, d_s, -1, n);
    hipLaunchKernelGGL(k_div_bad, dim3((uint32_t)n), dim3(64), 0, st, d_bad, qs, mu, rs, mv, n);
    return hipGetLastError();
}

// Grows the context's division buffers to what a batch of n instances of
// (mu, mv) digits at precision P needs, so that later calls up to that shape
// allocate (and free) nothing.
int div_reserve(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, uint32_t P) {
    int rc = set_device(ctx);
    if (rc) return rc;
    void* small;
    if ((rc = scratch(ctx, 4, (size_t)n * 16 + 256, &small))) return rc;
    uint32_t* snaps;
    uint64_t snap_words;
    if ((rc = mul_snaps(ctx, &snaps, &snap_words))) return rc;
    const DivLayout L = div_layout(mu, mv, P, n);
    DivState* ds = div_state(ctx);
    const size_t need = L.total_words * 4;
    if (ds->buf_bytes < need) {
        ds->drop_all();
        if (ds->buf) (void)hipFree(ds->buf);
        ds->buf = nullptr;
        ds->buf_bytes = 0;
        if (hipMalloc(&ds->buf, need) != hipSuccess) {
            (void)hipGetLastError();
            return BI_ERR_NOMEM;
        }
        ds->buf_bytes = need;
    }
    return BI_OK;
}

// graph: capture the pipeline in a hipGraph (cached per shape and buffers) and
// launch that; otherwise launch the same kernels directly on the stream --
// the server's division worker does this, because capturing and
// instantiating a graph for every new shape costs more than it saves there.
int divmod(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs, uint32_t* qs,
           uint32_t* rs, bool graph) {
    if (!ctx || mu == 0 || mv == 0) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    if (!us || !vs || !qs || !rs) return BI_ERR_ARG;
    // the long products stage both operands of an instance in one
    // workgroup's LDS (<= 160 KB): numerator + divisor within 16384 digits
    if (mu + std::max<uint32_t>(mv, 3) > 16384 || n > 0x7FFFFFFFull) return BI_ERR_SHAPE;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    // per-instance prep (outside the graph): significant lengths, shifts, max
    void* small;
    const size_t small_bytes = (size_t)n * 16 + 256;
    if ((rc = scratch(ctx, 4, small_bytes, &small))) return rc;
    if (divone_wanted(n, mu, mv, ctx->div_one_max)) {
        // small batch: the whole pipeline in one launch, one workgroup an
        // instance (bi_divone.hip), then the zero-divisor flags
        uint8_t* d_bad = (uint8_t*)small;
        hipError_t e = divone_launch(n, mu, mv, us, vs, qs, rs, d_bad, st);
        std::vector<uint8_t> bad(n);
        if (e == hipSuccess) e = hipMemcpyAsync(bad.data(), d_bad, n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(ctx, e);
        DivState* ds = div_state(ctx);
        ds->last_macs = ds->last_tmacs = 0;   // per-instance schedules: no batch figure
        ds->last_P = ds->last_steps = 0;
        for (uint64_t i = 0; i < n; i++)
            if (bad[i]) return BI_ERR_DIVZERO;
        return BI_OK;
    }
    uint32_t* sigu = (uint32_t*)small;
    uint32_t* sigv = sigu + n;
    int32_t* d_s = (int32_t*)(sigv + n);
    uint8_t* d_bad = (uint8_t*)(d_s + n);
    uint8_t* d_flag = d_bad + n;
    uint32_t* d_maxq = (uint32_t*)(((uintptr_t)(d_flag + n) + 15) & ~(uintptr_t)15);
    const uint32_t NV = std::max<uint32_t>(mv, 3);
    hipError_t e = topscan(2, us, nullptr, mu, n, sigu, st);
    if (e == hipSuccess) e = topscan(2, vs, nullptr, mv, n, sigv, st);
    if (e != hipSuccess) return hip_fail(ctx, e);
    hipLaunchKernelGGL(k_zero32, dim3(1), dim3(64), 0, st, d_maxq, (uint64_t)1);
    hipLaunchKernelGGL(k_div_prep, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, sigu, sigv, NV, d_s, d_bad,
                       d_maxq, n);
    uint32_t maxq = 0;
    e = hipMemcpyAsync(&maxq, d_maxq, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e);
    const uint32_t P = std::max<uint32_t>(maxq, 1) + 1;
    DivLayout L = div_layout(mu, mv, P, n);
    L.rows = (int)ctx->mul_cfg.rows;
    DivState* ds = div_state(ctx);
    const size_t need = L.total_words * 4;
    if (ds->buf_bytes < need) {
        // the cached graphs reference the old buffer
        ds->drop_all();
        if (ds->buf) (void)hipFree(ds->buf);
        ds->buf = nullptr;
        ds->buf_bytes = 0;
        e = hipMalloc(&ds->buf, need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return BI_ERR_NOMEM;
        }
        ds->buf_bytes = need;
    }
    // transform products among the Newton steps: their tables and stash are
    // set up here, outside any capture
    {
        uint32_t sched[64];
        const int ns = newton_schedule(L.PP, sched);
        bool any = rem_by_wrap(L.NV, L.NU, L.P, n);
        uint32_t p = 3;
        for (int i = 0; i < ns; i++) {
            any = any || vx_by_transform(L.NV, p + 1, n);
            p = sched[i];
        }
        if (any) {
            const uint64_t *wf, *wi;
            uint64_t* stash;
            if ((rc = ntt_prepare(ctx, n, &wf, &wi, &stash))) return rc;
        }
    }
    if (!graph) {
        BigEnv env = {};
        env.num_cus = (uint32_t)ctx->num_cus;
        if ((rc = mul_snaps(ctx, &env.snaps, &env.snap_words))) return rc;
        bigmul_prepare();
        double macs = 0, tmacs = 0;
        uint32_t steps = 0;
        e = div_record(ctx, L, (uint32_t*)ds->buf, us, vs, mu, mv, qs, rs, d_s, d_bad, d_flag, n, st, env, &macs,
                       &tmacs, &steps);
        if (e != hipSuccess) return hip_fail(ctx, e);
        ds->last_macs = macs * (double)n;
        ds->last_tmacs = tmacs * (double)n;
        ds->last_P = P;
        ds->last_steps = steps;
    } else {
        // the transform's tables and stash (scratch slots 10, 11) are part of
        // what a captured graph points at
        const uint64_t key[13] = {n, mu, mv, P, (uint64_t)us, (uint64_t)vs, (uint64_t)qs, (uint64_t)rs,
                                  (uint64_t)ds->buf, (uint64_t)small, (uint64_t)L.rows,
                                  (uint64_t)ctx->scratch[10], (uint64_t)ctx->scratch[11]};
        DivGraph* hit = nullptr;
        for (auto& x : ds->g)
            if (x.exec && std::equal(key, key + 13, x.key)) hit = &x;
        if (!hit) {
            DivGraph* slot = &ds->g[0];
            for (auto& x : ds->g)
                if (!x.exec || x.stamp < slot->stamp) {
                    slot = &x;
                    if (!x.exec) break;
                }
            ds->drop(*slot);
            bigmul_prepare();
            BigEnv env = {};
            env.num_cus = (uint32_t)ctx->num_cus;
            if ((rc = mul_snaps(ctx, &env.snaps, &env.snap_words))) return rc;
            e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(ctx, e);
            double macs = 0, tmacs = 0;
            uint32_t steps = 0;
            hipError_t er = div_record(ctx, L, (uint32_t*)ds->buf, us, vs, mu, mv, qs, rs, d_s, d_bad, d_flag, n, st, env,
                                       &macs, &tmacs, &steps);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(st, &g);
            if (er != hipSuccess || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(ctx, er != hipSuccess ? er : e);
            }
            e = hipGraphInstantiate(&slot->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                slot->exec = nullptr;
                (void)hipGraphDestroy(g);
                return hip_fail(ctx, e);
            }
            slot->graph = g;
            std::copy(key, key + 13, slot->key);
            slot->macs = macs * (double)n;
            slot->tmacs = tmacs * (double)n;
            slot->P = P;
            slot->steps = steps;
            hit = slot;
        }
        hit->stamp = ++ds->clock;
        ds->last_macs = hit->macs;
        ds->last_tmacs = hit->tmacs;
        ds->last_P = hit->P;
        ds->last_steps = hit->steps;
        e = hipGraphLaunch(hit->exec, st);
        if (e != hipSuccess) return hip_fail(ctx, e);
    }
    // report zero divisors (the other instances are complete)
    std::vector<uint8_t> bad(n);
    e = hipMemcpyAsync(bad.data(), d_bad, n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e);
    for (uint64_t i = 0; i < n; i++)
        if (bad[i]) return BI_ERR_DIVZERO;
    return BI_OK;
}

}  // namespace bi

using namespace bi;

extern "C" {

int bi_divmod_u32_dev(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs,
                      uint32_t* qs, uint32_t* rs) {
    return bi::divmod(ctx, n, mu, mv, us, vs, qs, rs, true);
}

int bi_tune_set_div_one(bi_ctx* ctx, uint32_t n_max) {
    if (!ctx) return BI_ERR_ARG;
    ctx->div_one_max = n_max;
    return BI_OK;
}

int bi_tune_get_div_one(bi_ctx* ctx, uint32_t* n_max) {
    if (!ctx || !n_max) return BI_ERR_ARG;
    *n_max = ctx->div_one_max;
    return BI_OK;
}

int bi_div_last_info(bi_ctx* ctx, uint32_t* P, uint32_t* newton_steps, double* macs) {
    if (!ctx) return BI_ERR_ARG;
    DivState* ds = static_cast<DivState*>(ctx->extras);
    if (!ds) return BI_ERR_STATE;
    if (P) *P = ds->last_P;
    if (newton_steps) *newton_steps = ds->last_steps;
    if (macs) *macs = ds->last_macs;
    return BI_OK;
}

int bi_div_last_work(bi_ctx* ctx, double* valu_macs, double* transform_macs) {
    if (!ctx) return BI_ERR_ARG;
    DivState* ds = static_cast<DivState*>(ctx->extras);
    if (!ds) return BI_ERR_STATE;
    if (valu_macs) *valu_macs = ds->last_macs - ds->last_tmacs;
    if (transform_macs) *transform_macs = ds->last_tmacs;
    return BI_OK;
}

int bi_divmod_u64_dev(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint64_t* us, const uint64_t* vs,
                      uint64_t* qs, uint64_t* rs) {
    if (mu > 0x7FFFFFFFu || mv > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return bi_divmod_u32_dev(ctx, n, 2 * mu, 2 * mv, (const uint32_t*)us, (const uint32_t*)vs, (uint32_t*)qs,
                             (uint32_t*)rs);
}

}  // extern "C"
