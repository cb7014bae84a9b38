// bi_serve.hip -- continuous-batching request server on one GPU.
//
// No counterpart in the reference (a thesis benchmark harness, cuda/main.cu).
//
// Requests (add, mul, div of integers of any size up to 4096 digits) arrive
// from host threads.  A scheduler thread admits waiting requests in arrival
// order while their device footprint fits the free part of a fixed HBM arena
// (admission by memory budget), packs the admitted add/mul requests into ONE
// ragged batch (add instances first, then mul instances, one CSR offset
// table) and launches it as one persistent kernel (bi_csr.hip) on one of
// several streams.  Division requests go to one of two division worker
// threads -- short (both operands <= kShortDivDigits) or long -- each with its
// own context, stream, arena region and pinned staging: each round it takes
// every division pending for it (up to kMaxDiv), zero-pads them to the longest
// operands and divides them in ONE launch (bi_divone.hip: a workgroup per
// division, each on its own precision schedule; the batched pipeline's
// buffers are reserved at start for shapes it would not fit).  A round costs
// about what its longest division costs (sequential Newton steps: 0.05 ms at
// 64 digits, 0.3 ms at 4096), so short divisions get rounds of their own
// rather than waiting for long ones.  Batches on different streams and the
// division rounds overlap on the device; requests that arrive while batches
// run are admitted as soon as arena space and a stream free up.  All data moves
// between pinned host staging and the arena through a staging kernel.  A
// request's latency is the time from its submission to the moment its result
// is in host memory.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/bigint.h"
#include "bi_ctx.h"

namespace {

using clk = std::chrono::steady_clock;

struct Req {
    uint64_t id;
    int op;                 // BI_OP_ADD, BI_OP_MUL, BI_OP_DIV
    uint32_t ma, mb;        // digits of a, b
    std::vector<uint32_t> a, b;
    std::vector<uint32_t> r0, r1;   // results: add/mul -> r0 (ma digits); div -> q (ma), r (mb)
    clk::time_point t_submit, t_done;
    int rc = 0;
    bool done = false;
};

// first-fit allocator over [0, size) in 256-byte units
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    void init(size_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t noff = it->first + bytes;
                free_.erase(it);
                if (rest) free_[noff] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (auto& kv : free_) s += kv.second;
        return s;
    }
};

constexpr uint32_t kMaxDiv = 256;          // divisions per worker round
constexpr uint32_t kMaxDivDigits = 4096;    // largest division operand
constexpr uint32_t kShortDivDigits = 512;   // largest operand of the short worker
constexpr int kDivWorkers = 2;              // [0] short, [1] long

// One division worker: its context (stream, division buffers and graph-free
// eager pipeline), its queue, its arena region and the pinned staging that
// mirrors the region.
struct DivWorker {
    bi_ctx* ctx = nullptr;
    std::thread th;
    std::deque<Req*> pending;
    uint32_t max_digits = 0;
    size_t off = 0, bytes = 0;
    uint32_t* host = nullptr;
    uint32_t* host_dev = nullptr;
};

struct Batch {
    std::vector<Req*> reqs;
    size_t arena_off = 0, arena_bytes = 0;
    int slot = -1;
    uint32_t* host = nullptr;    // the slot's pinned staging (results staged back here)
    size_t res_word_off = 0;     // where results start in the host staging
    uint64_t n_add = 0, n_mul = 0;
};

}  // namespace

struct bi_srv {
    bi_ctx* ctx = nullptr;       // the caller's context (arena allocation only)
    bi_ctx* sub[4] = {};         // one context (stream) per in-flight batch slot
    // per-slot resources reused batch after batch (a slot holds one batch at
    // a time and is reused only after that batch's stream has drained)
    bi_plan* slot_plan[4] = {};
    uint32_t* slot_host[4] = {};     // pinned staging, grow-only
    uint32_t* slot_host_dev[4] = {};  // its device address
    size_t slot_host_words[4] = {};
    int slots = 2;
    uint32_t max_batch = 4096;
    void* arena_base = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req*> waiting;
    std::map<uint64_t, Req*> all;
    std::vector<Batch*> inflight;
    std::atomic<bool> stop{false};
    std::atomic<bool> sched_done{false};
    std::thread th;
    // division workers: the Newton pipeline on their own contexts, so that
    // add/mul batches keep flowing while division rounds run; dw[1]'s context
    // also owns the arena
    DivWorker dw[kDivWorkers];
    std::condition_variable div_cv;

    uint64_t next_id = 1;
    int fatal = 0;
    // stats
    uint64_t batches = 0, max_batch_seen = 0;
    size_t peak_used = 0;
};

namespace {

size_t req_footprint(const Req* r) {
    // operands + results (+ CSR offset), padded
    return (size_t)r->ma * 3 * 4 + 64;
}

// arena region of a division worker whose operands reach D digits: up to
// kMaxDiv instances of D digits, at most arena/share (at least 4 instances)
size_t div_region_bytes(size_t arena, uint32_t D, size_t share) {
    return std::min((size_t)kMaxDiv * D * 4 * 4, std::max(arena / share, (size_t)4 * D * 16));
}

// operand lengths of a round: the longest in it, rounded up to a multiple of
// 4 digits (rows stay 16-byte aligned for the staging kernel)
uint32_t round4(uint32_t x) { return std::max<uint32_t>(4, (x + 3) & ~3u); }

// bytes of a round of `count` divisions with row lengths (mu, mv)
size_t div_round_bytes(uint32_t count, uint32_t mu, uint32_t mv) {
    return (size_t)count * (2 * (size_t)mu + 2 * (size_t)mv) * 4;
}

// Moves `words` digits between the pinned host staging and the arena, in
// either direction, with 16-byte loads (both pointers 16-byte aligned); as
// fast as the runtime's copy engines from 1 MiB up (about 55 GB/s), twice as
// fast below.  The server moves all its data with this kernel: copy-engine
// copies issued on several streams at once (the slots, the division worker)
// now and then stalled each other for about 6 ms, which set the tail latency.
__global__ void __launch_bounds__(256) k_stage(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                               uint64_t words) {
    const uint64_t nv = words / 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (uint64_t i = t; i < nv; i += stride) d4[i] = s4[i];
    for (uint64_t i = nv * 4 + t; i < words; i += stride) dst[i] = src[i];
}

int stage(bi_ctx* c, const uint32_t* src, uint32_t* dst, uint64_t words) {
    if (words == 0) return 0;
    if (((uintptr_t)src | (uintptr_t)dst) & 15) return BI_ERR_ARG;   // layouts keep both aligned
    const uint64_t blocks = std::min<uint64_t>(256, std::max<uint64_t>(1, (words / 4 + 255) / 256));
    hipLaunchKernelGGL(k_stage, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)bi_stream(c), src, dst, words);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : bi::hip_fail(c, e);
}

// Pinned host staging the device reads and writes directly: fine-grained
// (coherent), so no device cache holds a stale copy between batches.
int alloc_staging(bi_ctx* c, size_t bytes, uint32_t** host, uint32_t** dev) {
    *host = *dev = nullptr;
    hipError_t e = hipHostMalloc((void**)host, bytes, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)dev, *host, 0);
    if (e != hipSuccess) {
        if (*host) (void)hipHostFree(*host);
        *host = *dev = nullptr;
        (void)hipGetLastError();
        return e == hipErrorOutOfMemory ? BI_ERR_NOMEM : bi::hip_fail(c, e);
    }
    return 0;
}

// One division round: every request in `group`, zero-padded to the longest
// operands (rounded up to 4 digits).  Layout of the region and of the pinned
// staging: u | v | q | r.  (A round is one launch of the one-workgroup
// division, or the eager batched pipeline where its shape would not fit; no
// graph, so shapes need not repeat: no padding to size classes or counts.)
int run_div_round(bi_srv* s, DivWorker& w, std::vector<Req*>& group) {
    uint32_t mu = 4, mv = 4;
    for (Req* r : group) {
        mu = std::max(mu, round4(r->ma));
        mv = std::max(mv, round4(r->mb));
    }
    const uint32_t nc = (uint32_t)group.size();
    uint32_t* hu = w.host;
    uint32_t* hv = hu + (size_t)nc * mu;
    uint32_t* hq = hv + (size_t)nc * mv;
    uint32_t* hr = hq + (size_t)nc * mu;
    memset(hu, 0, ((size_t)nc * mu + (size_t)nc * mv) * 4);
    for (size_t i = 0; i < nc; i++) {
        memcpy(hu + i * mu, group[i]->a.data(), group[i]->ma * 4);
        memcpy(hv + i * mv, group[i]->b.data(), group[i]->mb * 4);
    }
    char* base = (char*)s->arena_base + w.off;
    uint32_t* du = (uint32_t*)base;
    uint32_t* dv = du + (size_t)nc * mu;
    uint32_t* dq = dv + (size_t)nc * mv;
    uint32_t* dr = dq + (size_t)nc * mu;
    hipStream_t st = (hipStream_t)bi_stream(w.ctx);
    const uint64_t words = (uint64_t)nc * mu + (uint64_t)nc * mv;
    uint32_t* hdev = w.host_dev;
    int rc = stage(w.ctx, hdev, du, words);
    if (rc) return rc;
    const int drc = bi::divmod(w.ctx, nc, mu, mv, du, dv, dq, dr, false);
    if (drc && drc != BI_ERR_DIVZERO) return drc;
    if ((rc = stage(w.ctx, dq, hdev + words, words))) return rc;
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return bi::hip_fail(w.ctx, e);
    // results copied before the lock, as in finish_batch
    for (size_t i = 0; i < group.size(); i++) {
        Req* r = group[i];
        r->r0.assign(hq + i * mu, hq + i * mu + r->ma);
        r->r1.assign(hr + i * mv, hr + i * mv + r->mb);
        bool zero = std::all_of(r->b.begin(), r->b.end(), [](uint32_t x) { return x == 0; });
        r->rc = zero ? BI_ERR_DIVZERO : 0;
    }
    const auto now = clk::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (Req* r : group) {
        r->t_done = now;
        r->done = true;
    }
    return 0;
}

// Launches the add/mul part of a batch asynchronously on its slot's stream.
int launch_addmul(bi_srv* s, Batch* b, std::vector<Req*>& adds, std::vector<Req*>& muls) {
    bi_ctx* c = s->sub[b->slot];
    const uint64_t n = adds.size() + muls.size();
    std::vector<int64_t> off(n + 1, 0);
    uint64_t k = 0;
    for (Req* r : adds) { off[k + 1] = off[k] + r->ma; k++; }
    for (Req* r : muls) { off[k + 1] = off[k] + r->ma; k++; }
    const uint64_t total = (uint64_t)off[n];
    b->n_add = adds.size();
    b->n_mul = muls.size();
    // device layout in the batch's arena chunk: off | a | b | r, the result
    // 16-byte aligned (the chunk's 4096 spare bytes cover the padding)
    const size_t offw = ((n + 1) * 2 + 3) & ~(size_t)3;   // offsets, in words
    const size_t rw = (2 * total + 3) & ~(uint64_t)3;      // a | b, padded
    char* base = (char*)s->arena_base + b->arena_off;
    int64_t* doff = (int64_t*)base;
    uint32_t* da = (uint32_t*)(base + ((offw * 4 + 255) & ~(size_t)255));
    uint32_t* db = da + total;
    uint32_t* dr = da + rw;
    // pinned staging: off | a | b | r, the same alignment
    const size_t need = offw + rw + total;
    if (s->slot_host_words[b->slot] < need) {
        const size_t words = std::max(need, s->slot_host_words[b->slot] * 2);
        if (s->slot_host[b->slot]) (void)hipHostFree(s->slot_host[b->slot]);
        s->slot_host[b->slot] = s->slot_host_dev[b->slot] = nullptr;
        s->slot_host_words[b->slot] = 0;
        const int arc = alloc_staging(c, words * 4, &s->slot_host[b->slot], &s->slot_host_dev[b->slot]);
        if (arc) return arc;
        s->slot_host_words[b->slot] = words;
    }
    b->host = s->slot_host[b->slot];
    memcpy(b->host, off.data(), (n + 1) * 8);
    uint32_t* ha = b->host + offw;
    uint32_t* hb = ha + total;
    k = 0;
    for (Req* r : adds) { memcpy(ha + off[k], r->a.data(), r->ma * 4); memcpy(hb + off[k], r->b.data(), r->ma * 4); k++; }
    for (Req* r : muls) { memcpy(ha + off[k], r->a.data(), r->ma * 4); memcpy(hb + off[k], r->b.data(), r->ma * 4); k++; }
    b->res_word_off = offw + rw;
    uint32_t* hdev = s->slot_host_dev[b->slot];
    int rc = stage(c, hdev, (uint32_t*)doff, (n + 1) * 2);
    if (!rc) rc = stage(c, hdev + offw, da, 2 * total);
    if (rc) return rc;
    if (!muls.empty()) {
        // the plan covers the mul instances only: sizes of the add instances
        // are zeroed in its size table (zero-digit instances are skipped)
        std::vector<int64_t> poff(n + 1);
        for (uint64_t i = 0; i <= n; i++) poff[i] = std::max<int64_t>(off[i], off[adds.size()]);
        rc = bi::csr_plan_rebuild(c, &s->slot_plan[b->slot], n, poff.data(), 0);
        if (rc) return rc;
    }
    bi_plan* plan = s->slot_plan[b->slot];
    if (!adds.empty() && !muls.empty()) {
        rc = bi::addmul_csr_split(c, plan, adds.size(), (uint64_t)off[adds.size()], doff, da, db, dr, dr);
    } else if (!adds.empty()) {
        rc = bi_add_csr_u32_dev(c, adds.size(), doff, (uint64_t)off[adds.size()], da, db, dr);
    } else {
        rc = bi_mul_csr_u32_dev(c, plan, doff, nullptr, da, db, dr);
    }
    if (rc) return rc;
    return stage(c, dr, hdev + b->res_word_off, total);
}

// Everything a first batch would otherwise pay for, done before the server
// takes requests: per slot the pinned staging (up to 64 MiB), plan room for
// max_batch instances, the multiply's snapshot scratch and one launch of each
// kernel a batch takes (staging, add only, multiply only, both); for each
// division worker three divisions of its largest class, so that every
// long-product kernel has run once on its context (its buffers are reserved
// at their largest shape before).  Without it the first batches stall on allocations and first
// launches while requests queue behind them.
int warm_up(bi_srv* s) {
    const size_t pin_words = std::min<size_t>(s->arena.size / (size_t)s->slots, (size_t)64 << 20) / 4;
    // two instances of 8 digits at the arena start: offsets | a | b | r
    char* base = (char*)s->arena_base;
    int64_t* doff = (int64_t*)base;
    uint32_t* da = (uint32_t*)(base + 256);
    uint32_t* db = da + 16;
    uint32_t* dr = db + 16;
    {
        std::vector<uint32_t> h(64 + 48, 0);
        const int64_t off[3] = {0, 8, 16};
        memcpy(h.data(), off, sizeof off);
        if (hipMemcpy(base, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            return BI_ERR_HIP;
        }
    }
    for (int k = 0; k < s->slots; k++) {
        bi_ctx* c = s->sub[k];
        int rc = alloc_staging(c, std::max<size_t>(pin_words, 4) * 4, &s->slot_host[k], &s->slot_host_dev[k]);
        if (rc) return rc;
        s->slot_host_words[k] = std::max<size_t>(pin_words, 4);
        rc = stage(c, s->slot_host_dev[k], da, 4);
        if (!rc) rc = bi::csr_plan_reserve(c, &s->slot_plan[k], s->max_batch);
        uint32_t* snaps;
        uint64_t snap_words;
        if (!rc) rc = bi::mul_snaps(c, &snaps, &snap_words);
        const int64_t poff[3] = {8, 8, 16};   // the add instance zeroed, as launch_addmul does
        if (!rc) rc = bi::csr_plan_rebuild(c, &s->slot_plan[k], 2, poff, 0);
        if (!rc) rc = bi::addmul_csr_split(c, s->slot_plan[k], 1, 8, doff, da, db, dr, dr);
        if (!rc) rc = bi_add_csr_u32_dev(c, 2, doff, 16, da, db, dr);
        if (!rc) rc = bi_mul_csr_u32_dev(c, s->slot_plan[k], doff, nullptr, da, db, dr);
        if (!rc) rc = bi_sync(c);
        if (rc) return rc;
    }
    // per worker, divisions of the largest class a round takes (numerator D
    // digits): over a D/2-digit divisor (precision D/2 + 1), over a 4-digit
    // one (the longest precision with the shortest divisor) and over a
    // D-digit one (the longest divisor); top digits 1, so the shapes are fixed
    for (DivWorker& w : s->dw) {
        uint32_t* du = (uint32_t*)((char*)s->arena_base + w.off);
        if (int rc = stage(w.ctx, w.host_dev, du, 4)) return rc;
        const uint32_t D = w.max_digits;
        const size_t nd = 4ull * D;
        std::vector<uint32_t> hu(nd, 0), hv(nd, 0);
        for (int i = 0; i < 4; i++) hu[(size_t)i * D + D - 1] = 1;
        uint32_t* dv = du + nd;
        uint32_t* dq = dv + nd;
        uint32_t* dd = dq + nd;
        if (div_round_bytes(4, D, D) > w.bytes || hipMemcpy(du, hu.data(), nd * 4, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            return BI_ERR_HIP;
        }
        for (const uint32_t mv : {D / 2, 4u, D}) {
            for (int i = 0; i < 4; i++) {
                std::fill(hv.begin() + (size_t)i * mv, hv.begin() + (size_t)(i + 1) * mv, 0u);
                hv[(size_t)i * mv] = 3;
                hv[(size_t)i * mv + mv - 1] = 1;
            }
            if (hipMemcpy(dv, hv.data(), 4ull * mv * 4, hipMemcpyHostToDevice) != hipSuccess) {
                (void)hipGetLastError();
                return BI_ERR_HIP;
            }
            if (int drc = bi::divmod(w.ctx, 4, D, mv, du, dv, dq, dd, false)) return drc;
        }
        if (int rc = bi_sync(w.ctx)) return rc;
    }
    return 0;
}

void div_worker(bi_srv* s, int wi) {
    DivWorker& w = s->dw[wi];
    (void)bi::set_device(w.ctx);
    for (;;) {
        std::vector<Req*> group;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->div_cv.wait(lk, [&] { return !w.pending.empty() || s->sched_done.load(); });
            if (w.pending.empty()) return;   // the scheduler has drained everything
            // FIFO, while the padded round fits the worker's region
            uint32_t mu = 4, mv = 4;
            while (!w.pending.empty() && group.size() < kMaxDiv) {
                const Req* r = w.pending.front();
                const uint32_t mu2 = std::max(mu, round4(r->ma)), mv2 = std::max(mv, round4(r->mb));
                if (div_round_bytes((uint32_t)group.size() + 1, mu2, mv2) > w.bytes) break;
                mu = mu2;
                mv = mv2;
                group.push_back(w.pending.front());
                w.pending.pop_front();
            }
        }
        const int rc = run_div_round(s, w, group);
        if (rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Req* r : group) {
                if (!r->done) {
                    r->rc = rc;
                    r->done = true;
                    r->t_done = clk::now();
                }
            }
            s->fatal = rc;
        }
        s->cv.notify_all();
    }
}

// Delivers the results of a batch whose stream has drained and releases its
// arena chunk.
void finish_batch(bi_srv* s, Batch* b) {
    const uint32_t* res = b->host ? b->host + b->res_word_off : nullptr;
    size_t o = 0;
    // add requests first, then mul, in the order they were packed; results
    // are copied before the lock (a request's result is read only after
    // `done`, which is set under it), so submitters never wait on the copies
    std::vector<Req*> order;
    for (Req* r : b->reqs) if (r->op == BI_OP_ADD) order.push_back(r);
    for (Req* r : b->reqs) if (r->op == BI_OP_MUL) order.push_back(r);
    if (res) {
        for (Req* r : order) {
            r->r0.assign(res + o, res + o + r->ma);
            o += r->ma;
        }
    }
    const auto now = clk::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (Req* r : order) {
        if (!res) break;
        r->t_done = now;
        r->done = true;
    }
    s->arena.release(b->arena_off, b->arena_bytes);
    delete b;
    s->cv.notify_all();
}

void scheduler(bi_srv* s) {
    (void)bi::set_device(s->dw[1].ctx);
    while (true) {
        // 1. retire finished batches
        for (size_t i = 0; i < s->inflight.size();) {
            Batch* b = s->inflight[i];
            hipStream_t st = (hipStream_t)bi_stream(s->sub[b->slot]);
            const hipError_t q = hipStreamQuery(st);
            if (q == hipSuccess) {
                s->inflight.erase(s->inflight.begin() + i);
                finish_batch(s, b);
            } else {
                if (q != hipErrorNotReady) s->fatal = BI_ERR_HIP;
                i++;
            }
        }
        // 2. admit a new batch into a free slot
        std::vector<Req*> take;
        Batch* nb = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop && s->waiting.empty() && s->inflight.empty() && s->dw[0].pending.empty() &&
                s->dw[1].pending.empty()) {
                s->sched_done = true;
                s->div_cv.notify_all();
                break;
            }
            int used = 0;
            for (Batch* b : s->inflight) used |= 1 << b->slot;
            int slot = -1;
            for (int k = 0; k < s->slots; k++)
                if (!(used & (1 << k))) { slot = k; break; }
            if (slot >= 0 && !s->waiting.empty()) {
                size_t foot = 4096;
                const size_t free_b = s->arena.free_bytes();
                while (!s->waiting.empty() && take.size() < s->max_batch) {
                    Req* r = s->waiting.front();
                    const size_t f = req_footprint(r);
                    if (foot + f > free_b) break;   // does not fit yet: wait for space
                    foot += f;
                    take.push_back(r);
                    s->waiting.pop_front();
                }
                if (!take.empty()) {
                    nb = new Batch();
                    nb->slot = slot;
                    nb->arena_bytes = foot;
                    if (!s->arena.alloc(foot, &nb->arena_off)) {
                        // fragmented: put the requests back and retry later
                        for (auto it = take.rbegin(); it != take.rend(); ++it) s->waiting.push_front(*it);
                        take.clear();
                        delete nb;
                        nb = nullptr;
                    } else {
                        const size_t used_b = s->arena.size - s->arena.free_bytes();
                        s->peak_used = std::max(s->peak_used, used_b);
                    }
                }
            }
            if (!nb && s->inflight.empty()) {
                s->cv.wait_for(lk, std::chrono::microseconds(200));
                continue;
            }
        }
        if (nb) {
            s->batches++;
            s->max_batch_seen = std::max<uint64_t>(s->max_batch_seen, take.size());
            std::vector<Req*> adds, muls;
            for (Req* r : take) {
                if (r->op == BI_OP_ADD) adds.push_back(r);
                else muls.push_back(r);
            }
            nb->reqs = adds;
            nb->reqs.insert(nb->reqs.end(), muls.begin(), muls.end());
            const int rc = launch_addmul(s, nb, adds, muls);
            if (rc) {
                // nothing of this batch will run: fail its requests now
                std::lock_guard<std::mutex> lk(s->mu);
                for (Req* r : take) {
                    r->rc = rc;
                    r->done = true;
                    r->t_done = clk::now();
                }
                nb->reqs.clear();
                s->fatal = rc;
            }
            s->inflight.push_back(nb);
        } else {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
}

// Frees whatever a server holds (null-safe: also for a start that failed
// half way), after its threads have ended.
void release_all(bi_srv* s) {
    for (int k = 0; k < s->slots; k++) {
        if (s->sub[k]) (void)bi_sync(s->sub[k]);
        bi::csr_plan_free(s->slot_plan[k]);
        if (s->slot_host[k]) (void)hipHostFree(s->slot_host[k]);
    }
    for (DivWorker& w : s->dw) {
        if (w.ctx) (void)bi_sync(w.ctx);
        if (w.host) (void)hipHostFree(w.host);
    }
    if (s->arena_base) (void)bi_dev_free(s->dw[1].ctx, s->arena_base);
    for (int k = 0; k < s->slots; k++)
        if (s->sub[k]) bi_destroy(s->sub[k]);
    for (DivWorker& w : s->dw)
        if (w.ctx) bi_destroy(w.ctx);
    delete s;
}

}  // namespace

extern "C" {

int bi_srv_start(bi_ctx* ctx, size_t arena_bytes, uint32_t max_batch, uint32_t slots, bi_srv** out) {
    if (!ctx || !out || arena_bytes < (1 << 20) || slots < 1 || slots > 4) return BI_ERR_ARG;
    *out = nullptr;
    bi_srv* s = new (std::nothrow) bi_srv();
    if (!s) return BI_ERR_NOMEM;
    s->ctx = ctx;
    s->slots = (int)slots;
    s->max_batch = max_batch ? max_batch : 4096;
    s->dw[0].max_digits = kShortDivDigits;
    s->dw[1].max_digits = kMaxDivDigits;
    int rc = 0;
    for (DivWorker& w : s->dw)
        if (!rc) rc = bi_create(&w.ctx, ctx->device);
    for (int k = 0; k < s->slots && !rc; k++) rc = bi_create(&s->sub[k], ctx->device);
    if (!rc) rc = bi_dev_alloc(s->dw[1].ctx, &s->arena_base, arena_bytes);
    // the division workers' buffers at their largest shapes: no allocation
    // (and no implicit device synchronisation of a free) while serving
    for (DivWorker& w : s->dw)
        if (!rc) rc = bi::div_reserve(w.ctx, kMaxDiv, w.max_digits, w.max_digits, 2 * w.max_digits);
    if (!rc) {
        // the division regions sit at the top of the arena (long, then short
        // below it); each worker's pinned staging mirrors its region
        DivWorker& lw = s->dw[1];
        DivWorker& sw = s->dw[0];
        lw.off = (arena_bytes - div_region_bytes(arena_bytes, lw.max_digits, 4)) & ~(size_t)255;
        lw.bytes = arena_bytes - lw.off;
        sw.off = (lw.off - div_region_bytes(arena_bytes, sw.max_digits, 16)) & ~(size_t)255;
        sw.bytes = lw.off - sw.off;
        s->arena.init(sw.off);
        for (DivWorker& w : s->dw)
            if (!rc) rc = alloc_staging(w.ctx, w.bytes, &w.host, &w.host_dev);
    }
    if (!rc) rc = warm_up(s);
    if (rc) {
        release_all(s);
        return rc;
    }
    s->th = std::thread(scheduler, s);
    for (int w = 0; w < kDivWorkers; w++) s->dw[w].th = std::thread(div_worker, s, w);
    *out = s;
    return BI_OK;
}

}  // extern "C"

namespace {

// bi_srv_submit with the request's arrival time given: latency is measured
// from it (the trace runner passes the scheduled arrival, so a late
// submission counts against the latency, as an open-loop client sees it).
int submit_at(bi_srv* s, int op, uint32_t ma, const uint32_t* a, uint32_t mb, const uint32_t* b, uint64_t* ticket,
              clk::time_point arrival) {
    if (!s || !ticket || !a || !b || ma == 0 || mb == 0) return BI_ERR_ARG;
    if (op != BI_OP_ADD && op != BI_OP_MUL && op != BI_OP_DIV) return BI_ERR_ARG;
    if (op != BI_OP_DIV && ma != mb) return BI_ERR_ARG;
    if (ma > 4096 || mb > 4096) return BI_ERR_SHAPE;
    if (s->stop) return BI_ERR_STATE;   // fast path; checked again under the lock
    Req* r = new (std::nothrow) Req();
    if (!r) return BI_ERR_NOMEM;
    r->op = op;
    r->ma = ma;
    r->mb = mb;
    r->a.assign(a, a + ma);
    r->b.assign(b, b + mb);
    if (op != BI_OP_DIV && req_footprint(r) + 4096 > s->arena.size) {
        delete r;
        return BI_ERR_NOMEM;
    }
    r->t_submit = arrival;
    std::lock_guard<std::mutex> lk(s->mu);
    // stop is set under this lock: a request enqueued here is seen by the
    // scheduler's (and the division workers') drain before they exit
    if (s->stop) {
        delete r;
        return BI_ERR_STATE;
    }
    r->id = s->next_id++;
    s->all[r->id] = r;
    *ticket = r->id;
    if (op == BI_OP_DIV) {
        s->dw[(ma <= kShortDivDigits && mb <= kShortDivDigits) ? 0 : 1].pending.push_back(r);
        s->div_cv.notify_all();
    } else {
        s->waiting.push_back(r);
        s->cv.notify_one();
    }
    return BI_OK;
}

}  // namespace

extern "C" {

int bi_srv_submit(bi_srv* s, int op, uint32_t ma, const uint32_t* a, uint32_t mb, const uint32_t* b,
                  uint64_t* ticket) {
    return submit_at(s, op, ma, a, mb, b, ticket, clk::now());
}

int bi_srv_poll(bi_srv* s, uint64_t ticket, uint32_t* r0, uint32_t* r1, int* done, double* latency_s) {
    if (!s || !done) return BI_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return BI_ERR_ARG;
    Req* r = it->second;
    *done = r->done ? 1 : 0;
    if (!r->done) return BI_OK;
    if (r0) std::copy(r->r0.begin(), r->r0.end(), r0);
    if (r1 && r->op == BI_OP_DIV) std::copy(r->r1.begin(), r->r1.end(), r1);
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    const int rc = r->rc;
    s->all.erase(it);
    delete r;
    return rc;
}

int bi_srv_stats(bi_srv* s, uint64_t* batches, uint64_t* max_batch, size_t* peak_arena_bytes) {
    if (!s) return BI_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches) *batches = s->batches;
    if (max_batch) *max_batch = s->max_batch_seen;
    if (peak_arena_bytes) *peak_arena_bytes = s->peak_used;
    return s->fatal;
}

int bi_srv_stop(bi_srv* s) {
    if (!s) return BI_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->div_cv.notify_all();
    }
    for (DivWorker& w : s->dw)
        if (w.th.joinable()) w.th.join();
    const int fatal = s->fatal;
    for (auto& kv : s->all) delete kv.second;
    release_all(s);
    return fatal;
}

}  // extern "C"

extern "C" {

// Open-loop driver: submits request i at arrival_s[i] seconds after the call
// starts (sleeping in between), then collects every result.  Operands of
// request i are a_pool[a_off[i] ..) and b_pool[b_off[i] ..); results go to
// r0_pool[r_off0[i] ..) (and r1_pool[r_off1[i] ..) for divisions).
int bi_srv_run_trace(bi_srv* s, uint64_t n, const int32_t* ops, const uint32_t* ma, const uint32_t* mb,
                     const uint64_t* a_off, const uint32_t* a_pool, const uint64_t* b_off, const uint32_t* b_pool,
                     const double* arrival_s, uint32_t* r0_pool, const uint64_t* r_off0, uint32_t* r1_pool,
                     const uint64_t* r_off1, double* latency_s, int32_t* status, double* wall_s) {
    if (!s || (n && (!ops || !ma || !mb || !a_off || !a_pool || !b_off || !b_pool || !arrival_s || !r0_pool ||
                     !r_off0 || !latency_s || !status)))
        return BI_ERR_ARG;
    std::vector<uint64_t> tickets(n, 0);
    const auto t0 = clk::now();
    for (uint64_t i = 0; i < n; i++) {
        const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(arrival_s[i]));
        // keep the open-loop schedule: an OS sleep overshoots by tens of us,
        // more than the gap between arrivals at these rates, so only the
        // part of a wait beyond 200 us sleeps and the rest spins
        for (;;) {
            const auto now = clk::now();
            if (now >= due) break;
            if (due - now > std::chrono::microseconds(200))
                std::this_thread::sleep_for(due - now - std::chrono::microseconds(150));
            else
                std::this_thread::yield();
        }
        int rc;
        while ((rc = submit_at(s, ops[i], ma[i], a_pool + a_off[i], mb[i], b_pool + b_off[i], &tickets[i], due)) ==
               BI_ERR_BUSY)
            std::this_thread::sleep_for(std::chrono::microseconds(10));
        if (rc) return rc;
    }
    // completion times are stamped by the scheduler, so collecting results in
    // submission order does not change any latency
    const auto deadline = clk::now() + std::chrono::seconds(300);
    for (uint64_t i = 0; i < n; i++) {
        for (;;) {
            if (clk::now() > deadline) return BI_ERR_STATE;
            int done = 0;
            const int rc = bi_srv_poll(s, tickets[i], r0_pool + r_off0[i],
                                       (ops[i] == BI_OP_DIV && r1_pool) ? r1_pool + r_off1[i] : nullptr, &done,
                                       &latency_s[i]);
            if (done) {
                status[i] = rc;
                break;
            }
            if (rc) return rc;
            const int f = bi_srv_stats(s, nullptr, nullptr, nullptr);
            if (f) return f;
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
    if (wall_s) *wall_s = std::chrono::duration<double>(clk::now() - t0).count();
    return BI_OK;
}

}  // extern "C"
