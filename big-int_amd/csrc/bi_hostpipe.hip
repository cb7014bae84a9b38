// bi_hostpipe.hip -- the host-pointer entry points' transfer pipeline.
//
// The reference's gpuAdd / gpuMultiply copy both operand arrays in, run the
// kernel and copy the result out, one step after the other (cuda/main.cu:
// 121-128,214-221).  Here a batch larger than one chunk moves through pinned
// staging in chunks of whole instances on two streams: while chunk k's
// kernel runs, chunk k+1's operands go in and chunk k-1's result comes out.
// The copies between the caller's (pageable) arrays and the pinned staging
// are host memcpy on a small team of threads -- one thread moves about
// 10 GB/s, the link several times that.  Kernels of consecutive chunks are
// ordered by events: they share the context's scratch (chunk states,
// multiply snapshots).
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/bigint.h"
#include "bi_ctx.h"

namespace bi {

namespace {

constexpr size_t kChunkBytes = 8u << 20;      // per operand array and chunk
constexpr size_t kPipeMinBytes = 4u << 20;    // smaller batches: one plain round trip
constexpr unsigned kCopyThreads = 8;          // host copy team (caller included)

// A persistent team of host threads for large memcpy: run() splits one copy
// into equal parts, the caller does part 0, the workers the rest.
class CopyTeam {
  public:
    explicit CopyTeam(unsigned nthreads) {
        for (unsigned i = 1; i < nthreads; i++) th_.emplace_back([this, i] { work(i); });
        n_ = nthreads;
    }
    ~CopyTeam() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void run(void* dst, const void* src, size_t bytes) {
        if (bytes < (1u << 20) || n_ == 1) {
            memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            pending_ = n_ - 1;
            gen_++;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
    }

  private:
    void part(unsigned i) {
        const size_t per = ((bytes_ + n_ - 1) / n_ + 63) & ~(size_t)63;
        const size_t lo = std::min(bytes_, per * i), hi = std::min(bytes_, lo + per);
        if (hi > lo) memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void work(unsigned i) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
                if (quit_) return;
                seen = gen_;
            }
            part(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    unsigned n_ = 1;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    unsigned pending_ = 0;
    uint64_t gen_ = 0;
    bool quit_ = false;
};

}  // namespace

struct HostPipe {
    hipStream_t st[2] = {};
    hipEvent_t kern_done[2] = {}, out_done[2] = {};
    uint32_t* pin = nullptr;    // 2 slots x (a | b | r) of chunk words each
    size_t pin_words = 0;
    CopyTeam* team = nullptr;
};

void hostpipe_free(bi_ctx* ctx) {
    HostPipe* p = static_cast<HostPipe*>(ctx->hostpipe);
    if (!p) return;
    for (int s = 0; s < 2; s++) {
        if (p->st[s] && p->st[s] != ctx->stream) {
            (void)hipStreamSynchronize(p->st[s]);
            (void)hipStreamDestroy(p->st[s]);
        }
        if (p->kern_done[s]) (void)hipEventDestroy(p->kern_done[s]);
        if (p->out_done[s]) (void)hipEventDestroy(p->out_done[s]);
    }
    if (p->pin) (void)hipHostFree(p->pin);
    delete p->team;
    delete p;
    ctx->hostpipe = nullptr;
}

static int hostpipe_get(bi_ctx* ctx, size_t chunk_words, HostPipe** out) {
    HostPipe* p = static_cast<HostPipe*>(ctx->hostpipe);
    if (!p) {
        p = new (std::nothrow) HostPipe();
        if (!p) return BI_ERR_NOMEM;
        ctx->hostpipe = p;
        p->st[0] = ctx->stream;
        hipError_t e = hipStreamCreateWithFlags(&p->st[1], hipStreamNonBlocking);
        for (int s = 0; s < 2 && e == hipSuccess; s++) {
            e = hipEventCreateWithFlags(&p->kern_done[s], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&p->out_done[s], hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            hostpipe_free(ctx);
            return hip_fail(ctx, e);
        }
        const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
        p->team = new (std::nothrow) CopyTeam(std::min(kCopyThreads, hc));
        if (!p->team) {
            hostpipe_free(ctx);
            return BI_ERR_NOMEM;
        }
    }
    if (p->pin_words < 6 * chunk_words) {
        if (p->pin) {
            for (int s = 0; s < 2; s++) (void)hipStreamSynchronize(p->st[s]);
            (void)hipHostFree(p->pin);
            p->pin = nullptr;
            p->pin_words = 0;
        }
        void* h = nullptr;
        if (hipHostMalloc(&h, 6 * chunk_words * 4, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return BI_ERR_NOMEM;
        }
        p->pin = static_cast<uint32_t*>(h);
        p->pin_words = 6 * chunk_words;
    }
    *out = p;
    return BI_OK;
}

// rs = op(as, bs) for n instances of m digits in host memory.  launch(st, a,
// b, r, count) enqueues the kernels of `count` instances on stream st.
int host_pipelined(bi_ctx* ctx, uint32_t m, uint64_t n, const uint32_t* as, const uint32_t* bs, uint32_t* rs,
                   const HostLaunch& launch, uint64_t chunk_hint) {
    const size_t row = (size_t)m * 4;
    uint64_t c = chunk_hint ? chunk_hint : std::max<uint64_t>(1, kChunkBytes / row);
    if ((uint64_t)n * row < kPipeMinBytes || c >= n) c = n;
    const size_t cw = (size_t)c * m;    // words per array and chunk
    HostPipe* p;
    int rc = hostpipe_get(ctx, cw, &p);
    if (rc) return rc;
    void *da, *db, *dr;
    if ((rc = scratch(ctx, 0, 2 * cw * 4, &da)) || (rc = scratch(ctx, 1, 2 * cw * 4, &db)) ||
        (rc = scratch(ctx, 2, 2 * cw * 4, &dr)))
        return rc;
    // the second stream starts after whatever the context stream holds
    hipError_t e0 = hipEventRecord(p->kern_done[0], p->st[0]);
    if (e0 == hipSuccess) e0 = hipStreamWaitEvent(p->st[1], p->kern_done[0], 0);
    if (e0 != hipSuccess) return hip_fail(ctx, e0);
    const uint64_t chunks = (n + c - 1) / c;
    auto count_of = [&](uint64_t k) { return std::min<uint64_t>(c, n - k * c); };
    auto pin_of = [&](int s, int which) { return p->pin + ((size_t)s * 3 + which) * cw; };
    auto dev_of = [&](void* base, int s) { return static_cast<uint32_t*>(base) + (size_t)s * cw; };
    auto copy_out = [&](uint64_t k) -> int {
        const int s = (int)(k & 1);
        hipError_t e = hipEventSynchronize(p->out_done[s]);
        if (e != hipSuccess) return hip_fail(ctx, e);
        p->team->run(rs + k * c * m, pin_of(s, 2), count_of(k) * row);
        return BI_OK;
    };
    for (uint64_t k = 0; k < chunks; k++) {
        const int s = (int)(k & 1);
        const size_t bytes = count_of(k) * row;
        if (k >= 2 && (rc = copy_out(k - 2))) return rc;
        p->team->run(pin_of(s, 0), as + k * c * m, bytes);
        p->team->run(pin_of(s, 1), bs + k * c * m, bytes);
        hipStream_t st = p->st[s];
        hipError_t e = hipMemcpyAsync(dev_of(da, s), pin_of(s, 0), bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dev_of(db, s), pin_of(s, 1), bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess && k >= 1) e = hipStreamWaitEvent(st, p->kern_done[s ^ 1], 0);
        if (e != hipSuccess) return hip_fail(ctx, e);
        e = launch(st, dev_of(da, s), dev_of(db, s), dev_of(dr, s), count_of(k));
        if (e == hipSuccess) e = hipEventRecord(p->kern_done[s], st);
        if (e == hipSuccess) e = hipMemcpyAsync(pin_of(s, 2), dev_of(dr, s), bytes, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(p->out_done[s], st);
        if (e != hipSuccess) return hip_fail(ctx, e);
    }
    for (uint64_t k = chunks >= 2 ? chunks - 2 : 0; k < chunks; k++)
        if ((rc = copy_out(k))) return rc;
    return BI_OK;
}

}  // namespace bi
