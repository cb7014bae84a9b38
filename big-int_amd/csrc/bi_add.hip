// bi_add.hip -- batched add / subtract kernels (uniform and ragged batches).
// The span algorithm is in bi_addcore.h.
#include <algorithm>

#include "bi_addcore.h"
#include "bi_kernels.h"

namespace bi {

// One wave per span of instances [w*ipw, min((w+1)*ipw, n)); 4 waves per
// workgroup (uniform: size m each; CSR: as the offset table says).
template <bool SUB, bool CSR, bool VEC, bool NT, int U, bool CMP = false>
__global__ void __launch_bounds__(256) k_addsub(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                uint32_t* __restrict__ rs, uint64_t n, uint32_t m, uint64_t ipw,
                                                const int64_t* __restrict__ off, uint64_t nspans,
                                                uint8_t* __restrict__ order) {
    const uint32_t wv = threadIdx.x / kWave;
    const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x / kWave) + wv;
    if (w >= nspans) return;
    addsub_span<SUB, CSR, VEC, NT, U, CMP>(as, bs, rs, n, m, ipw, off, w, order);
}

// One wave per chunk of D digits (cpi chunks per instance); lane 0 records
// the chunk's carry state (add_chunk) in st[w].
template <bool VEC, bool NT, int U>
__global__ void __launch_bounds__(256) k_add_chunks(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                    uint32_t* __restrict__ rs, uint32_t m, uint32_t D, uint32_t cpi,
                                                    uint64_t nchunks, uint8_t* __restrict__ st) {
    const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (w >= nchunks) return;
    const uint32_t state = add_chunk<VEC, NT, U>(as, bs, rs, m, D, cpi, w);
    if (lane_id() == 0) st[w] = (uint8_t)state;
}

// Carries between the chunks of each instance (one wave per instance): the
// chunk states are the generate / propagate bits of the same wave look-ahead
// as the digits (bi::wave_carry), 64 chunks a step.  A chunk whose carry-in
// is 1 gets +1 from its first digit up to its first digit that is not
// all-ones, which lies inside the chunk unless the chunk is all-ones (then
// every digit becomes zero and the carry goes on, as the look-ahead says).
__global__ void __launch_bounds__(256) k_chunk_carry(uint32_t* __restrict__ rs, uint32_t m, uint32_t D, uint32_t cpi,
                                                     uint64_t n, const uint8_t* __restrict__ st) {
    const uint32_t lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (i >= n) return;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < cpi; c0 += kWave) {
        const uint32_t c = c0 + lane;
        const uint32_t s = c < cpi ? st[i * cpi + c] : 0u;
        uint32_t cout;
        const uint32_t e = wave_carry(s == 1u, s == 2u, carry, cout);
        carry = cout;
        if (c < cpi && e) {
            uint32_t* r = rs + i * m + (uint64_t)c * D;
            const uint32_t len = (uint64_t)c * D + D < m ? D : m - c * D;
            if (s == 2u) {
                for (uint32_t k = 0; k < len; k++) r[k] = 0u;
            } else {
                for (uint32_t k = 0; k < len; k++) {
                    const uint32_t x = r[k] + 1u;
                    r[k] = x;
                    if (x) break;
                }
            }
        }
    }
}

// One wave per D-digit chunk of a ragged batch's flat digit array
// (add_flat_chunk).  Chunks in blockIdx order: handing each XCD a contiguous
// eighth of them (so a chunk's look-back would hit the L2 its neighbour just
// filled) measured 104.3 us against 101.7 us for 2^17 instances of 8..2048
// digits.
template <bool NT, int U>
__global__ void __launch_bounds__(256) k_add_flat(const uint32_t* __restrict__ as, const uint32_t* __restrict__ bs,
                                                  uint32_t* __restrict__ rs, const int64_t* __restrict__ off,
                                                  uint64_t n, uint64_t total, uint32_t D, uint64_t nchunks) {
    const uint64_t w = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (w >= nchunks) return;
    add_flat_chunk<NT, U>(as, bs, rs, off, n, total, D, w);
}

// Chunks of a uniform add: D = span digits rounded down to whole tiles, used
// when an instance is longer than one chunk.
static uint32_t chunk_digits(const AddCfg& cfg, uint32_t m) {
    const uint32_t D = cfg.span_digits & ~255u;
    return (D >= 256 && m > D) ? D : 0u;
}

uint64_t add_chunk_states(const AddCfg& cfg, uint32_t m, uint64_t n) {
    const uint32_t D = chunk_digits(cfg, m);
    return D ? n * ((m + D - 1) / D) : 0;
}

template <bool SUB, bool CSR>
static hipError_t launch_addsub(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint64_t n,
                                uint32_t m, const int64_t* off, uint64_t total, bool vec, hipStream_t st,
                                uint8_t* order = nullptr) {
    uint64_t ipw = 1;
    if (!CSR) {
        ipw = cfg.span_digits / m;
    } else {
        // whole instances, about span_digits digits a span, at most 63 (one
        // offset per lane)
        ipw = total ? (cfg.span_digits * n + total / 2) / total : 63;
        if (ipw > 63) ipw = 63;
    }
    if (ipw < 1) ipw = 1;
    const uint64_t nspans = (n + ipw - 1) / ipw;
    if (nspans == 0) return hipSuccess;
    // |a - b|: with at most 63 instances a span, each wave compares its own
    // instances (one launch); otherwise a compare pass writes `order` first
    const bool cmp = SUB && !CSR && order && ipw <= 63;
    if (SUB && !CSR && order && !cmp) {
        const hipError_t e = topscan(1, as, bs, m, n, order, st);
        if (e != hipSuccess) return e;
    }
    const uint32_t wpb = 4;
    const dim3 grid((uint32_t)((nspans + wpb - 1) / wpb)), block(wpb * kWave);
#define BI_LAUNCH(VEC, NT, U)                                                                                      \
    do {                                                                                                           \
        if (cmp)                                                                                                   \
            hipLaunchKernelGGL((k_addsub<SUB, CSR, VEC, NT, U, SUB && !CSR>), grid, block, 0, st, as, bs, rs, n, m,  \
                               ipw, off, nspans, order);                                                           \
        else                                                                                                       \
            hipLaunchKernelGGL((k_addsub<SUB, CSR, VEC, NT, U>), grid, block, 0, st, as, bs, rs, n, m, ipw, off,     \
                               nspans, order);                                                                     \
    } while (0)
    if (vec) {
        if (cfg.nontemporal) {
            if (cfg.tiles_in_flight >= 4) BI_LAUNCH(true, true, 4);
            else if (cfg.tiles_in_flight >= 2) BI_LAUNCH(true, true, 2);
            else BI_LAUNCH(true, true, 1);
        } else {
            if (cfg.tiles_in_flight >= 4) BI_LAUNCH(true, false, 4);
            else if (cfg.tiles_in_flight >= 2) BI_LAUNCH(true, false, 2);
            else BI_LAUNCH(true, false, 1);
        }
    } else {
        BI_LAUNCH(false, false, 1);
    }
#undef BI_LAUNCH
    return hipGetLastError();
}

hipError_t add_uniform(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                       uint64_t n, bool sub, hipStream_t st, uint8_t* order, uint8_t* chunk_states) {
    const bool vec = ((((uintptr_t)as) | ((uintptr_t)bs) | ((uintptr_t)rs)) & 15u) == 0;
    const uint32_t D = chunk_digits(cfg, m);
    if (!sub && D && chunk_states) {
        const uint32_t cpi = (m + D - 1) / D;
        const uint64_t nchunks = n * cpi;
        const dim3 grid((uint32_t)((nchunks + 3) / 4)), block(256);
#define BI_CHUNKS(VEC, NT, U)                                                                                    \
    hipLaunchKernelGGL((k_add_chunks<VEC, NT, U>), grid, block, 0, st, as, bs, rs, m, D, cpi, nchunks, chunk_states)
        if (!vec) BI_CHUNKS(false, false, 1);
        else if (cfg.nontemporal) {
            if (cfg.tiles_in_flight >= 4) BI_CHUNKS(true, true, 4);
            else if (cfg.tiles_in_flight >= 2) BI_CHUNKS(true, true, 2);
            else BI_CHUNKS(true, true, 1);
        } else {
            if (cfg.tiles_in_flight >= 4) BI_CHUNKS(true, false, 4);
            else if (cfg.tiles_in_flight >= 2) BI_CHUNKS(true, false, 2);
            else BI_CHUNKS(true, false, 1);
        }
#undef BI_CHUNKS
        hipLaunchKernelGGL(k_chunk_carry, dim3((uint32_t)((n + 3) / 4)), dim3(256), 0, st, rs, m, D, cpi, n,
                           chunk_states);
        return hipGetLastError();
    }
    return sub ? launch_addsub<true, false>(cfg, as, bs, rs, n, m, nullptr, n * m, vec, st, order)
               : launch_addsub<false, false>(cfg, as, bs, rs, n, m, nullptr, n * m, vec, st);
}

hipError_t add_csr(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, const int64_t* off,
                   uint64_t n, uint64_t total, bool sub, hipStream_t st) {
    const bool vec = ((((uintptr_t)as) | ((uintptr_t)bs) | ((uintptr_t)rs)) & 15u) == 0;
    if (!sub && vec && n && total) {
        // fixed chunks of the flat array: span digits rounded down to whole tiles
        const uint32_t D = std::max<uint32_t>(256, cfg.span_digits & ~255u);
        const uint64_t nchunks = (total + D - 1) / D;
        const dim3 grid((uint32_t)((nchunks + 3) / 4)), block(256);
#define BI_FLAT(NT, U) hipLaunchKernelGGL((k_add_flat<NT, U>), grid, block, 0, st, as, bs, rs, off, n, total, D, nchunks)
        if (cfg.nontemporal) {
            if (cfg.tiles_in_flight >= 8) BI_FLAT(true, 8);
            else if (cfg.tiles_in_flight >= 4) BI_FLAT(true, 4);
            else if (cfg.tiles_in_flight >= 2) BI_FLAT(true, 2);
            else BI_FLAT(true, 1);
        } else {
            if (cfg.tiles_in_flight >= 8) BI_FLAT(false, 8);
            else if (cfg.tiles_in_flight >= 4) BI_FLAT(false, 4);
            else if (cfg.tiles_in_flight >= 2) BI_FLAT(false, 2);
            else BI_FLAT(false, 1);
        }
#undef BI_FLAT
        return hipGetLastError();
    }
    return sub ? launch_addsub<true, true>(cfg, as, bs, rs, n, 0, off, total, vec, st)
               : launch_addsub<false, true>(cfg, as, bs, rs, n, 0, off, total, vec, st);
}

}  // namespace bi
