// bi_csr.hip -- ragged (CSR) batches: integers of different digit counts
// stored back to back, instance i in digits [off[i], off[i+1]).
//
// The reference has no ragged path: its batches are [n][m] with one m per
// launch (cuda/main.cu:131-135,303-307), so a mixed batch must be padded to a
// size class and launched once per class.  Here one launch covers the whole
// batch:
//   add -- the flat segmented carry scan of bi_add.hip, spans balanced by
//          digit count (no plan needed);
//   mul -- a plan packs every instance into a "team" of exactly the lanes its
//          size needs (ceil(M4/8) for the low product, M4/4 for the full one),
//          teams sorted by size and packed into workgroup "bundles" so that a
//          wave holds teams of similar step counts; teams never straddle a
//          wave unless they own whole waves.  A persistent grid pulls bundles
//          (largest first) and add spans from one work counter.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/bigint.h"
#include "bi_ctx.h"
#include "bi_kernels.h"
#include "bi_mulcore.h"

namespace bi {

struct CsrTeam {
    uint32_t inst;      // instance index
    uint32_t lane_off;  // first lane in the bundle
    uint32_t T;         // lanes
    uint32_t lds_off;   // first LDS word
};
struct CsrBundle {
    uint32_t team_begin, team_end;
};

// Team LDS layout: META[4] | U with the operand / column layout of
// bi_mulcore.h (KW = M8 for the low product, 2*M8 for the full one); the
// first-group sums recorded at a lane's switch go to global scratch.
__host__ __device__ inline uint32_t csr_team_lanes(uint32_t m, bool full) {
    const uint32_t M8 = (m + 7) & ~7u;
    return full ? M8 / 4 : M8 / 8;
}
__host__ __device__ inline uint32_t csr_team_words(uint32_t m, bool full) {
    const uint32_t M8 = (m + 7) & ~7u;
    return team_words(M8, full ? 2 * M8 : M8);
}

// One workgroup runs one bundle.
template <bool FULL, int R>
__device__ __forceinline__ void run_bundle(const CsrBundle bd, const CsrTeam* __restrict__ teams,
                                           const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                           uint32_t* __restrict__ rs, const int64_t* __restrict__ off,
                                           const int64_t* __restrict__ off_out, uint32_t* lds, uint32_t* snap) {
    const uint32_t tid = threadIdx.x;
    // team of this thread: last team with lane_off <= tid
    uint32_t lo = bd.team_begin, hi = bd.team_end;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (teams[mid].lane_off <= tid) lo = mid;
        else hi = mid;
    }
    const CsrTeam tm = teams[lo];
    const uint32_t tl = tid - tm.lane_off;
    const bool member = tid >= tm.lane_off && tl < tm.T;
    const int64_t o = off[tm.inst];
    const uint32_t m = (uint32_t)(off[tm.inst + 1] - o);
    const uint32_t M8 = (m + 7) & ~7u;
    const uint32_t KW = FULL ? 2 * M8 : M8;
    const uint32_t CW = team_col_stride(KW);
    uint32_t* META = lds + tm.lds_off;
    uint32_t* U = META + 4;
    uint32_t* A = U;
    uint32_t* W = U + M8 + 12;
    uint32_t* E = U;
    uint32_t* O = U + CW;

    if (member) {
        for (uint32_t i = 4 * tl; i < kWinPad; i += 4 * tm.T) *reinterpret_cast<u32x4*>(W + i) = u32x4{0u, 0u, 0u, 0u};
        if (tl == 0) {
            META[0] = 0;
            META[1] = 0;
        }
    }
    __syncthreads();
    if (member) {
        const uint32_t* ga = as + o;
        const uint32_t* gb = bs + o;
        uint32_t pa = 0, pb = 0;
        for (uint32_t i = 4 * tl; i < M8 + 16; i += 4 * tm.T) {
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
    __syncthreads();

    LaneSched sc = {0, 0, 0, 0, 0, 0};
    uint32_t K = 0;
    if (member) {
        const uint32_t na = META[0], nb = META[1];
        const uint32_t Nf = ((na > nb ? na : nb) + 7) & ~7u;
        if (FULL || 2 * Nf <= M8) {
            sc = sched_rect_r<R>(tl, Nf / 4, Nf / 4);
            K = 2 * Nf;
        } else {
            sc = sched_low_r<R>(tl, M8 / 4);
            K = M8;
        }
    }
    const uint32_t S = readfirst(wave_max_u32(sc.stepsA + sc.stepsB));
    uint64_t acc[4];
    uint32_t cnt[4];
    conv_walk<R>(A, W + 16 - 2 * R, sc, S, snap, acc, cnt);
    __syncthreads();   // every lane is done with the operands
    if (member) {
        put_lane_groups(E, O, sc, snap, acc, cnt);
        zero_groups(E, O, K, KW, tl, tm.T);
    }
    __syncthreads();

    // ---- recombination: E[k] + O[k] (< 2^33) split into a low word
    // and a carry bit added one digit up
    const uint32_t out_len = FULL ? 2 * m : m;
    uint32_t* out = rs + (FULL ? off_out[tm.inst] : o);
    if (tm.T <= 64) {
        // 8 digits per lane, one segmented carry look-ahead per wave
        uint32_t s[8];
        uint32_t c = 0, all = 0xFFFFFFFFu;
        const uint32_t k0 = 8 * tl;
        const bool any = member && k0 < out_len;
        if (any) {
            const u32x4 e0 = *reinterpret_cast<const u32x4*>(E + k0);
            const u32x4 e1 = *reinterpret_cast<const u32x4*>(E + k0 + 4);
            const u32x4 o0 = *reinterpret_cast<const u32x4*>(O + k0);
            const u32x4 o1 = *reinterpret_cast<const u32x4*>(O + k0 + 4);
            const uint32_t km = k0 ? k0 - 1 : 0;
            uint32_t hp;
            (void)__builtin_addc(E[km], O[km], 0u, &hp);
            hp = k0 ? hp : 0u;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                uint32_t h;
                const uint32_t x = j < 4 ? __builtin_addc(e0[j], o0[j], 0u, &h) : __builtin_addc(e1[j - 4], o1[j - 4], 0u, &h);
                s[j] = __builtin_addc(x, hp, c, &c);
                all &= s[j];
                hp = h;
            }
        }
        const bool start = member && tl == 0;
        const bool g = any && c != 0;
        const bool p = any && !start && all == 0xFFFFFFFFu;
        uint32_t cout;
        uint32_t e = wave_carry(g, p, 0u, cout);
        if (start || !any) e = 0;
        if (any) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint32_t k = k0 + j;
                if (k < out_len) out[k] = s[j] + e;
                e &= (uint32_t)(s[j] == 0xFFFFFFFFu);
            }
        }
    } else if (member && tl < 64) {
        // the team owns whole waves; its first wave walks 256-digit tiles
        const uint32_t lane = tl;
        uint32_t carry = 0;
        for (uint32_t q0 = 0; q0 < out_len; q0 += 256) {
            const uint32_t k = q0 + 4 * lane;
            uint32_t s4[4] = {0u, 0u, 0u, 0u}, c = 0, all = 0xFFFFFFFFu;
            if (k < out_len) {
                const u32x4 e4 = *reinterpret_cast<const u32x4*>(E + k);
                const u32x4 o4 = *reinterpret_cast<const u32x4*>(O + k);
                const uint32_t km = k ? k - 1 : 0;
                uint32_t hp;
                (void)__builtin_addc(E[km], O[km], 0u, &hp);
                hp = k ? hp : 0u;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    uint32_t h;
                    const uint32_t x = __builtin_addc(e4[j], o4[j], 0u, &h);
                    s4[j] = __builtin_addc(x, hp, c, &c);
                    all &= s4[j];
                    hp = h;
                }
            } else {
                all = 0;
            }
            uint32_t cout;
            uint32_t e = wave_carry(c != 0, all == 0xFFFFFFFFu, carry, cout);
            carry = cout;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (k + j < out_len) out[k + j] = s4[j] + e;
                e &= (uint32_t)(s4[j] == 0xFFFFFFFFu);
            }
        }
    }
}

// Work items 0..nb-1 are multiply bundles; items nb..nb+nspans-1 are add spans
// (the digit stream of the same batch, carried by wave 0 of the workgroup
// through the segmented carry scan).
struct CsrLaunch {
    const uint32_t* as;
    const uint32_t* bs;
    uint32_t* rs_mul;
    uint32_t* rs_add;
    const int64_t* off;
    const int64_t* off_out;
    const CsrTeam* teams;
    const CsrBundle* bundles;
    uint32_t nbundles;
    uint64_t n;
    uint64_t add_ipw;    // instances per add span, <= 63 (0: no add work)
    uint64_t nspans;
    uint32_t* work;      // persistent work counter (zeroed before the launch)
    uint32_t* snaps;     // per-lane snapshot scratch (12 words per lane of the grid)
};

__device__ void add_span_wave(const CsrLaunch& L, uint64_t w);

template <bool FULL, int R>
__global__ void __launch_bounds__(1024) k_csr(CsrLaunch L) {
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    __shared__ uint32_t item_s;
    const uint32_t wpb = blockDim.x / kWave;
    const uint64_t total = (uint64_t)L.nbundles + (L.add_ipw ? (L.nspans + wpb - 1) / wpb : 0);
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(L.work, 1u);
        __syncthreads();
        const uint32_t item = item_s;
        __syncthreads();
        if (item >= total) break;
        if (item < L.nbundles) {
            run_bundle<FULL, R>(L.bundles[item], L.teams, L.as, L.bs, L.rs_mul, L.off, L.off_out, lds,
                                L.snaps + ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 12);
        } else {
            // one add span per wave
            const uint32_t wv = threadIdx.x / kWave;
            const uint64_t w = (uint64_t)(item - L.nbundles) * wpb + wv;
            if (w < L.nspans) add_span_wave(L, w);
        }
        __syncthreads();
    }
}

}  // namespace bi

#include "bi_addcore.h"

namespace bi {

__device__ void add_span_wave(const CsrLaunch& L, uint64_t w) {
    const bool vec = ((((uintptr_t)L.as) | ((uintptr_t)L.bs) | ((uintptr_t)L.rs_add)) & 15u) == 0;
    if (vec) addsub_span<false, true, true, true, 2>(L.as, L.bs, L.rs_add, L.n, 0, L.add_ipw, L.off, w, nullptr);
    else addsub_span<false, true, false, false, 1>(L.as, L.bs, L.rs_add, L.n, 0, L.add_ipw, L.off, w, nullptr);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace bi

// ----------------------------------------------------------------- plans
struct bi_plan {
    int mode;                   // 0 low n_i digits, 1 full 2 n_i product
    uint64_t n;
    uint64_t total_digits;
    uint32_t nbundles, block, lds_bytes;
    uint32_t max_digits;
    bi::CsrTeam* d_teams = nullptr;
    bi::CsrBundle* d_bundles = nullptr;
    uint32_t* d_work = nullptr;
    int device = 0;
    double useful_macs = 0;     // sum n_i(n_i+1)/2 (low) or n_i^2 (full)
    size_t teams_cap = 0, bundles_cap = 0;      // device capacities (reuse)
    std::vector<bi::CsrTeam> h_teams;           // host copies (live while an
    std::vector<bi::CsrBundle> h_bundles;       //  asynchronous upload runs)
};

using namespace bi;

namespace {

// Packs instances (largest first) into bundles of <= lane_budget lanes and
// <= lds_budget words; a team of <= 64 lanes stays inside one wave.
void pack(const std::vector<uint32_t>& sizes, bool full, uint32_t lane_budget, uint32_t lds_budget,
          std::vector<CsrTeam>& teams, std::vector<CsrBundle>& bundles, uint32_t& block, uint32_t& lds_words) {
    std::vector<uint32_t> order(sizes.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return sizes[x] > sizes[y]; });
    uint32_t lanes = 0, words = 0, begin = 0;
    block = 64;
    lds_words = 0;
    auto flush = [&]() {
        if (teams.size() > begin) {
            bundles.push_back({begin, (uint32_t)teams.size()});
            block = std::max(block, (lanes + 63) / 64 * 64);
            lds_words = std::max(lds_words, words);
        }
        begin = (uint32_t)teams.size();
        lanes = 0;
        words = 0;
    };
    for (uint32_t i : order) {
        const uint32_t m = sizes[i];
        if (m == 0) continue;
        const uint32_t T = csr_team_lanes(m, full);
        const uint32_t W = (csr_team_words(m, full) + 3) & ~3u;
        for (int pass = 0; pass < 2; pass++) {
            uint32_t start, need;
            if (T > 64) {
                start = (lanes + 63) / 64 * 64;
                need = (T + 63) / 64 * 64;
            } else {
                start = ((lanes % 64) + T > 64) ? (lanes + 63) / 64 * 64 : lanes;
                need = T;
            }
            const bool fits = start + need <= std::max(lane_budget, need) && words + W <= std::max(lds_budget, W);
            if (!fits && teams.size() > begin && pass == 0) {
                flush();
                continue;
            }
            teams.push_back({i, start, T, words});
            lanes = start + need;
            words += W;
            break;
        }
    }
    flush();
}

}  // namespace

namespace {

// Packs a plan into p, reusing its device arrays when they are large enough.
// async: upload on the context stream (ordered before later launches on it).
int plan_build(bi_ctx* ctx, uint64_t n, const int64_t* offs, int mode, uint32_t lane_budget, bi_plan* p,
               bool async) {
    if (n > 0xFFFFFFFFull) return BI_ERR_SHAPE;
    std::vector<uint32_t> sizes(n);
    uint32_t maxd = 0;
    double useful = 0;
    for (uint64_t i = 0; i < n; i++) {
        const int64_t d = offs[i + 1] - offs[i];
        if (d < 0) return BI_ERR_ARG;
        if (d > (mode ? 2048 : 4096)) return BI_ERR_SHAPE;
        sizes[i] = (uint32_t)d;
        maxd = std::max(maxd, (uint32_t)d);
        useful += mode ? (double)d * d : (double)d * (d + 1) / 2;
    }
    if (lane_budget == 0) lane_budget = ctx->csr_lanes ? ctx->csr_lanes : 256;
    if (lane_budget < 64 || lane_budget > 1024 || (lane_budget & 63)) return BI_ERR_ARG;
    p->h_teams.clear();
    p->h_bundles.clear();
    uint32_t block = 64, words = 0;
    // LDS budget: about 24 KB per 256 lanes.  One launch gives every
    // workgroup the largest bundle's LDS, so a bundle of many small teams
    // sets the occupancy of all of them; 24 KB keeps 5-6 workgroups a CU
    // (40 KB allowed 3: the 2^17-instance batch of 8..2048 digits took 8 %
    // longer, 16 KB packs small teams too thinly).
    const uint32_t lds_budget = 6144 * lane_budget / 256;
    pack(sizes, mode == 1, lane_budget, lds_budget, p->h_teams, p->h_bundles, block, words);
    if ((size_t)words * 4 > 159 * 1024 || block > 1024) return BI_ERR_SHAPE;
    p->mode = mode;
    p->n = n;
    p->total_digits = n ? (uint64_t)offs[n] : 0;
    p->nbundles = (uint32_t)p->h_bundles.size();
    p->block = block;
    p->lds_bytes = words * 4;
    p->max_digits = maxd;
    p->device = ctx->device;
    p->useful_macs = useful;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    if (p->teams_cap < p->h_teams.size() || !p->d_teams) {
        if (p->d_teams) (void)hipFree(p->d_teams);
        p->teams_cap = std::max<size_t>(1, p->h_teams.size());
        e = hipMalloc(&p->d_teams, p->teams_cap * sizeof(CsrTeam));
    }
    if (e == hipSuccess && (p->bundles_cap < p->h_bundles.size() || !p->d_bundles)) {
        if (p->d_bundles) (void)hipFree(p->d_bundles);
        p->bundles_cap = std::max<size_t>(1, p->h_bundles.size());
        e = hipMalloc(&p->d_bundles, p->bundles_cap * sizeof(CsrBundle));
    }
    if (e == hipSuccess && !p->d_work) e = hipMalloc(&p->d_work, 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BI_ERR_NOMEM;
    }
    if (async) {
        if (!p->h_teams.empty())
            e = hipMemcpyAsync(p->d_teams, p->h_teams.data(), p->h_teams.size() * sizeof(CsrTeam),
                               hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess && !p->h_bundles.empty())
            e = hipMemcpyAsync(p->d_bundles, p->h_bundles.data(), p->h_bundles.size() * sizeof(CsrBundle),
                               hipMemcpyHostToDevice, ctx->stream);
    } else {
        if (!p->h_teams.empty())
            e = hipMemcpy(p->d_teams, p->h_teams.data(), p->h_teams.size() * sizeof(CsrTeam), hipMemcpyHostToDevice);
        if (e == hipSuccess && !p->h_bundles.empty())
            e = hipMemcpy(p->d_bundles, p->h_bundles.data(), p->h_bundles.size() * sizeof(CsrBundle),
                          hipMemcpyHostToDevice);
    }
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

}  // namespace

namespace bi {

// (Re)builds *pp for a new batch on ctx's stream without any device-wide
// synchronisation: the previous use of the plan must have completed on that
// stream (the caller retires batches in stream order).
int csr_plan_rebuild(bi_ctx* ctx, bi_plan** pp, uint64_t n, const int64_t* offs, int mode) {
    if (!*pp) {
        *pp = new (std::nothrow) bi_plan();
        if (!*pp) return BI_ERR_NOMEM;
    }
    return plan_build(ctx, n, offs, mode, 0, *pp, true);
}

// Gives *pp device room for `teams` teams and bundles, so that rebuilds of
// batches up to that many instances allocate (and free) nothing.
int csr_plan_reserve(bi_ctx* ctx, bi_plan** pp, size_t teams) {
    if (!*pp) {
        *pp = new (std::nothrow) bi_plan();
        if (!*pp) return BI_ERR_NOMEM;
    }
    bi_plan* p = *pp;
    int rc = set_device(ctx);
    if (rc) return rc;
    teams = std::max<size_t>(1, teams);
    hipError_t e = hipSuccess;
    if (p->teams_cap < teams) {
        if (p->d_teams) (void)hipFree(p->d_teams);
        p->d_teams = nullptr;
        p->teams_cap = 0;
        e = hipMalloc(&p->d_teams, teams * sizeof(CsrTeam));
        if (e == hipSuccess) p->teams_cap = teams;
    }
    if (e == hipSuccess && p->bundles_cap < teams) {
        if (p->d_bundles) (void)hipFree(p->d_bundles);
        p->d_bundles = nullptr;
        p->bundles_cap = 0;
        e = hipMalloc(&p->d_bundles, teams * sizeof(CsrBundle));
        if (e == hipSuccess) p->bundles_cap = teams;
    }
    if (e == hipSuccess && !p->d_work) e = hipMalloc(&p->d_work, 256);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return BI_ERR_NOMEM;
    }
    p->h_teams.reserve(teams);
    p->h_bundles.reserve(teams);
    return BI_OK;
}

void csr_plan_free(bi_plan* p) {
    if (!p) return;
    if (p->d_teams) (void)hipFree(p->d_teams);
    if (p->d_bundles) (void)hipFree(p->d_bundles);
    if (p->d_work) (void)hipFree(p->d_work);
    delete p;
}

}  // namespace bi

extern "C" {

int bi_csr_plan_create(bi_ctx* ctx, uint64_t n, const int64_t* offs, int mode, uint32_t lane_budget,
                       bi_plan** out) {
    if (!ctx || !out || (n && !offs) || (mode != 0 && mode != 1)) return BI_ERR_ARG;
    *out = nullptr;
    bi_plan* p = new (std::nothrow) bi_plan();
    if (!p) return BI_ERR_NOMEM;
    const int rc = plan_build(ctx, n, offs, mode, lane_budget, p, false);
    if (rc) {
        bi::csr_plan_free(p);
        return rc;
    }
    *out = p;
    return BI_OK;
}

int bi_csr_plan_destroy(bi_plan* p) {
    if (!p) return BI_ERR_ARG;
    (void)hipSetDevice(p->device);
    (void)hipDeviceSynchronize();
    if (p->d_teams) (void)hipFree(p->d_teams);
    if (p->d_bundles) (void)hipFree(p->d_bundles);
    if (p->d_work) (void)hipFree(p->d_work);
    delete p;
    return BI_OK;
}

int bi_csr_plan_info(const bi_plan* p, uint32_t* nbundles, uint32_t* block, uint32_t* lds_bytes, double* useful_macs) {
    if (!p) return BI_ERR_ARG;
    if (nbundles) *nbundles = p->nbundles;
    if (block) *block = p->block;
    if (lds_bytes) *lds_bytes = p->lds_bytes;
    if (useful_macs) *useful_macs = p->useful_macs;
    return BI_OK;
}

// add spans cover instances [0, n_add) (add_total digits); multiply bundles
// come from the plan.
static int csr_launch(bi_ctx* ctx, const bi_plan* plan, uint64_t n_add, uint64_t add_total, const int64_t* offs,
                      const int64_t* offs_out, const uint32_t* as, const uint32_t* bs, uint32_t* rs_add,
                      uint32_t* rs_mul) {
    int rc = set_device(ctx);
    if (rc) return rc;
    CsrLaunch L;
    L.as = as;
    L.bs = bs;
    L.rs_mul = rs_mul;
    L.rs_add = rs_add;
    L.off = offs;
    L.off_out = offs_out;
    L.teams = plan ? plan->d_teams : nullptr;
    L.bundles = plan ? plan->d_bundles : nullptr;
    L.nbundles = (plan && rs_mul) ? plan->nbundles : 0;
    L.n = n_add;
    // add spans of whole instances, about 4096 digits each (at most 63
    // instances: one offset per lane); the standalone ragged add's tuned
    // chunk is a different kernel's shape
    L.add_ipw = 0;
    if (rs_add && add_total && n_add) {
        const uint64_t digits = 4096;
        L.add_ipw = std::min<uint64_t>(63, std::max<uint64_t>(1, (digits * n_add + add_total / 2) / add_total));
    }
    L.nspans = L.add_ipw ? (n_add + L.add_ipw - 1) / L.add_ipw : 0;
    uint32_t* work = plan ? plan->d_work : nullptr;
    L.work = work;
    const uint32_t block = plan ? std::max<uint32_t>(plan->block, 256) : 256;
    const uint32_t wpb = block / kWave;
    const uint64_t items = (uint64_t)L.nbundles + (L.add_ipw ? (L.nspans + wpb - 1) / wpb : 0);
    if (items == 0) return BI_OK;
    const size_t lds = (plan && rs_mul) ? plan->lds_bytes : 0;
    hipError_t e;
    // persistent grid: as many workgroups as fit on the device at once
    const uint32_t per_cu_lds = lds ? (uint32_t)std::max<size_t>(1, (160 * 1024) / (lds + 1024)) : 8;
    const uint32_t per_cu_thr = std::max<uint32_t>(1, 2048 / block);
    const uint32_t per_cu = std::min(per_cu_lds, per_cu_thr);
    uint64_t grid = (uint64_t)ctx->num_cus * per_cu;
    if (grid > items) grid = items;
    uint64_t snap_words = 0;
    if ((rc = mul_snaps(ctx, &L.snaps, &snap_words))) return rc;
    if (grid * block * 12 > snap_words) grid = snap_words / (12ull * block);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, work, 1u);
    // once per process, thread-safe (static initialisation); static LDS of
    // k_csr: the work item word (4 B)
    static const bool attr = [] {
        set_max_lds(reinterpret_cast<const void*>(&k_csr<false, 4>), 160 * 1024 - 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_csr<true, 4>), 160 * 1024 - 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_csr<false, 8>), 160 * 1024 - 1024);
        set_max_lds(reinterpret_cast<const void*>(&k_csr<true, 8>), 160 * 1024 - 1024);
        return true;
    }();
    (void)attr;
    const dim3 g3((uint32_t)grid), b3(block);
    const bool full = plan && plan->mode == 1;
    if (ctx->csr_rows == 4) {
        if (full) hipLaunchKernelGGL((k_csr<true, 4>), g3, b3, lds, ctx->stream, L);
        else hipLaunchKernelGGL((k_csr<false, 4>), g3, b3, lds, ctx->stream, L);
    } else {
        if (full) hipLaunchKernelGGL((k_csr<true, 8>), g3, b3, lds, ctx->stream, L);
        else hipLaunchKernelGGL((k_csr<false, 8>), g3, b3, lds, ctx->stream, L);
    }
    e = hipGetLastError();
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_add_csr_u32_dev(bi_ctx* ctx, uint64_t n, const int64_t* offs, uint64_t total_digits, const uint32_t* as,
                       const uint32_t* bs, uint32_t* rs) {
    if (!ctx || (n && (!offs || !as || !bs || !rs))) return BI_ERR_ARG;
    if (n == 0 || total_digits == 0) return BI_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    hipError_t e = add_csr(ctx->csr_add_cfg, as, bs, rs, offs, n, total_digits, false, ctx->stream);
    return e == hipSuccess ? BI_OK : hip_fail(ctx, e);
}

int bi_mul_csr_u32_dev(bi_ctx* ctx, const bi_plan* plan, const int64_t* offs, const int64_t* offs_out,
                       const uint32_t* as, const uint32_t* bs, uint32_t* rs) {
    if (!ctx || !plan || (plan->n && (!offs || !as || !bs || !rs))) return BI_ERR_ARG;
    if (plan->mode == 1 && !offs_out) return BI_ERR_ARG;
    if (plan->device != ctx->device) return BI_ERR_ARG;
    return csr_launch(ctx, plan, 0, 0, offs, offs_out, as, bs, nullptr, rs);
}

int bi_addmul_csr_u32_dev(bi_ctx* ctx, const bi_plan* plan, const int64_t* offs, const int64_t* offs_out,
                          const uint32_t* as, const uint32_t* bs, uint32_t* rs_add, uint32_t* rs_mul) {
    if (!ctx || !plan || (plan->n && (!offs || !as || !bs || !rs_add || !rs_mul))) return BI_ERR_ARG;
    if (plan->mode == 1 && !offs_out) return BI_ERR_ARG;
    if (plan->device != ctx->device) return BI_ERR_ARG;
    return csr_launch(ctx, plan, plan->n, plan->total_digits, offs, offs_out, as, bs, rs_add, rs_mul);
}

}  // extern "C"

namespace bi {

// add for instances [0, n_add) and the plan's multiplies, one launch
int addmul_csr_split(bi_ctx* ctx, const bi_plan* plan, uint64_t n_add, uint64_t add_total, const int64_t* offs,
                     const uint32_t* as, const uint32_t* bs, uint32_t* rs_add, uint32_t* rs_mul) {
    return csr_launch(ctx, plan, n_add, add_total, offs, nullptr, as, bs, rs_add, rs_mul);
}

}  // namespace bi
