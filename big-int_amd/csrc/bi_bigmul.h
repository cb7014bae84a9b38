// bi_bigmul.h -- long products and windowed add/sub (internal).
#pragma once
#include "bi_device.h"

namespace bi {

// A window of a per-instance digit buffer: digits [off, off+len) of
// p + inst*stride.
struct Win {
    const uint32_t* p;
    uint64_t stride;
    uint32_t off;
    uint32_t len;
};

// Launch environment of the long products: column-sum scratch (cs, at least
// bigmul_scratch_words(a.len, b.len, n) words), stream, rows per block step
// (4 or 8), and the per-lane snapshot scratch of the persistent triangle
// kernel (num_cus * 2048 lanes * 12 words).
struct BigEnv {
    uint32_t* cs;
    size_t cs_words;
    hipStream_t st;
    int rows;
    uint32_t* snaps;
    uint64_t snap_words;
    uint32_t num_cus;
};

// out[inst*out_stride + k - lo] = digit k of (a*b) for k in [lo, lo+len);
// neg (lo == 0 only): the digits of (B^len - a*b) mod B^len instead.
hipError_t bigmul(Win a, Win b, uint32_t* out, uint64_t out_stride, uint32_t lo, uint32_t len, uint64_t n,
                  const BigEnv& env, bool neg = false);
// digits [0, K) of a*b (triangle of the low columns), or of -a*b mod B^K
hipError_t bigmul_low(Win a, Win b, uint32_t* out, uint64_t out_stride, uint32_t K, uint64_t n, const BigEnv& env,
                      bool neg = false);
// digits [lo, lo+len) of a*b from the top columns only (columns < lo-2
// dropped): the exact window or one less
hipError_t bigmul_high(Win a, Win b, uint32_t* out, uint64_t out_stride, uint32_t lo, uint32_t len, uint64_t n,
                       const BigEnv& env);
// sets kernel attributes; call once before any stream capture
void bigmul_prepare();
size_t bigmul_scratch_words(uint32_t alen, uint32_t blen, uint64_t n);

// r = X*B^dx +/- Y*B^dy (+ cin), len digits per instance (mod B^len);
// pred[i] == 0 (when pred is given) makes r = X*B^dx for that instance.
hipError_t win_addsub(bool sub, Win x, int32_t dx, Win y, int32_t dy, uint32_t* r, uint64_t r_stride, uint32_t len,
                      uint32_t cin, const uint8_t* pred, uint64_t n, hipStream_t st);

}  // namespace bi
