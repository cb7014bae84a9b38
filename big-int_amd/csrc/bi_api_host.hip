// bi_api_host.hip -- host-pointer forms of the division and of the ragged
// (CSR) add and multiply: the caller's arrays stay in host memory and each
// call is one round trip (device buffers for the call, operands in, the
// device entry point, results out), as the reference's own host entry points
// are (cuda/main.cu:121-128,214-221: allocate, copy in, launch, copy out,
// free; prototypes/div.c:427-467 divides arrays in host memory).  Output
// digits an entry point does not write (before offs[0], between instances
// of a full product) keep the caller's values: the output array is copied in
// as well.
#include <stdint.h>

#include <vector>

#include "../../include/bigint.h"

namespace {

// Device buffers of one call, freed on every return path.
struct CallBufs {
    bi_ctx* ctx;
    std::vector<void*> p;
    explicit CallBufs(bi_ctx* c) : ctx(c) {}
    ~CallBufs() {
        for (void* x : p) (void)bi_dev_free(ctx, x);
    }
    int get(void** out, size_t bytes) {
        const int rc = bi_dev_alloc(ctx, out, bytes);
        if (rc == BI_OK) p.push_back(*out);
        return rc;
    }
};

// n * digits * word bytes without overflow
bool bytes_of(uint64_t n, uint64_t digits, uint64_t word, size_t* out) {
    if (digits && n > SIZE_MAX / word / digits) return false;
    *out = (size_t)(n * digits * word);
    return true;
}

// Offsets of a ragged batch (host): n + 1 entries, offs[0] >= 0,
// non-decreasing; *total = offs[n].
bool offsets_ok(uint64_t n, const int64_t* offs, uint64_t* total) {
    if (offs[0] < 0) return false;
    for (uint64_t i = 0; i < n; i++)
        if (offs[i + 1] < offs[i]) return false;
    *total = (uint64_t)offs[n];
    return true;
}

}  // namespace

extern "C" {

int bi_divmod_u32(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint32_t* us, const uint32_t* vs,
                  uint32_t* qs, uint32_t* rs) {
    if (!ctx || (n && (!us || !vs || !qs || !rs))) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    size_t bu, bv;
    if (!bytes_of(n, mu, 4, &bu) || !bytes_of(n, mv, 4, &bv)) return BI_ERR_SHAPE;
    CallBufs B(ctx);
    void *du, *dv, *dq, *dr;
    int rc;
    if ((rc = B.get(&du, bu)) || (rc = B.get(&dv, bv)) || (rc = B.get(&dq, bu)) || (rc = B.get(&dr, bv))) return rc;
    if ((rc = bi_copy_h2d(ctx, du, us, bu)) || (rc = bi_copy_h2d(ctx, dv, vs, bv))) return rc;
    const int drc = bi_divmod_u32_dev(ctx, n, mu, mv, static_cast<const uint32_t*>(du),
                                      static_cast<const uint32_t*>(dv), static_cast<uint32_t*>(dq),
                                      static_cast<uint32_t*>(dr));
    if (drc != BI_OK && drc != BI_ERR_DIVZERO) return drc;   // zero divisors: every other instance is complete
    if ((rc = bi_copy_d2h(ctx, qs, dq, bu)) || (rc = bi_copy_d2h(ctx, rs, dr, bv))) return rc;
    return drc;
}

int bi_divmod_u64(bi_ctx* ctx, uint64_t n, uint32_t mu, uint32_t mv, const uint64_t* us, const uint64_t* vs,
                  uint64_t* qs, uint64_t* rs) {
    // a u64 limb is two u32 digits, low first (little-endian words)
    if (mu > 0x7FFFFFFFu || mv > 0x7FFFFFFFu) return BI_ERR_SHAPE;
    return bi_divmod_u32(ctx, n, 2 * mu, 2 * mv, reinterpret_cast<const uint32_t*>(us),
                         reinterpret_cast<const uint32_t*>(vs), reinterpret_cast<uint32_t*>(qs),
                         reinterpret_cast<uint32_t*>(rs));
}

int bi_add_csr_u32(bi_ctx* ctx, uint64_t n, const int64_t* offs, const uint32_t* as, const uint32_t* bs,
                   uint32_t* rs) {
    if (!ctx || (n && (!offs || !as || !bs || !rs))) return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    uint64_t total;
    if (!offsets_ok(n, offs, &total)) return BI_ERR_ARG;
    if (total == 0) return BI_OK;
    size_t bd, bo;
    if (!bytes_of(1, total, 4, &bd) || !bytes_of(n + 1, 1, 8, &bo)) return BI_ERR_SHAPE;
    CallBufs B(ctx);
    void *doff, *da, *db, *dr;
    int rc;
    if ((rc = B.get(&doff, bo)) || (rc = B.get(&da, bd)) || (rc = B.get(&db, bd)) || (rc = B.get(&dr, bd))) return rc;
    if ((rc = bi_copy_h2d(ctx, doff, offs, bo)) || (rc = bi_copy_h2d(ctx, da, as, bd)) ||
        (rc = bi_copy_h2d(ctx, db, bs, bd)) || (rc = bi_copy_h2d(ctx, dr, rs, bd)))
        return rc;
    if ((rc = bi_add_csr_u32_dev(ctx, n, static_cast<const int64_t*>(doff), total, static_cast<const uint32_t*>(da),
                                 static_cast<const uint32_t*>(db), static_cast<uint32_t*>(dr))))
        return rc;
    return bi_copy_d2h(ctx, rs, dr, bd);
}

int bi_mul_csr_u32(bi_ctx* ctx, uint64_t n, const int64_t* offs, const int64_t* offs_out, const uint32_t* as,
                   const uint32_t* bs, uint32_t* rs, int mode) {
    if (!ctx || (mode != 0 && mode != 1) || (n && (!offs || !as || !bs || !rs)) || (mode == 1 && n && !offs_out))
        return BI_ERR_ARG;
    if (n == 0) return BI_OK;
    uint64_t total, rtotal;
    if (!offsets_ok(n, offs, &total)) return BI_ERR_ARG;
    if (mode == 1) {
        // every instance's 2 n_i digits fit its output window
        if (!offsets_ok(n, offs_out, &rtotal)) return BI_ERR_ARG;
        for (uint64_t i = 0; i < n; i++)
            if (offs_out[i + 1] - offs_out[i] < 2 * (offs[i + 1] - offs[i])) return BI_ERR_ARG;
    } else {
        rtotal = total;
    }
    if (total == 0) return BI_OK;
    size_t bd, br, bo;
    if (!bytes_of(1, total, 4, &bd) || !bytes_of(1, rtotal, 4, &br) || !bytes_of(n + 1, 1, 8, &bo))
        return BI_ERR_SHAPE;
    bi_plan* plan = nullptr;
    int rc = bi_csr_plan_create(ctx, n, offs, mode, 0, &plan);
    if (rc) return rc;
    {
        CallBufs B(ctx);
        void *doff, *doo = nullptr, *da, *db, *dr;
        if ((rc = B.get(&doff, bo)) || (mode == 1 && (rc = B.get(&doo, bo))) || (rc = B.get(&da, bd)) ||
            (rc = B.get(&db, bd)) || (rc = B.get(&dr, br)))
            goto done;
        if ((rc = bi_copy_h2d(ctx, doff, offs, bo)) || (mode == 1 && (rc = bi_copy_h2d(ctx, doo, offs_out, bo))) ||
            (rc = bi_copy_h2d(ctx, da, as, bd)) || (rc = bi_copy_h2d(ctx, db, bs, bd)) ||
            (rc = bi_copy_h2d(ctx, dr, rs, br)))
            goto done;
        if ((rc = bi_mul_csr_u32_dev(ctx, plan, static_cast<const int64_t*>(doff), static_cast<const int64_t*>(doo),
                                     static_cast<const uint32_t*>(da), static_cast<const uint32_t*>(db),
                                     static_cast<uint32_t*>(dr))))
            goto done;
        rc = bi_copy_d2h(ctx, rs, dr, br);
    done:;
    }
    const int prc = bi_csr_plan_destroy(plan);
    return rc ? rc : prc;
}

}  // extern "C"
