// bi_kernels.h -- launchers of the gfx950 kernels (internal to libbigint).
#pragma once
#include "bi_device.h"

namespace bi {

// Raises a kernel's dynamic LDS limit (for workgroups above 64 KB).  `bytes`
// must leave room for the kernel's static LDS; a failure is cleared so it
// cannot surface as the error of a later launch.
inline void set_max_lds(const void* fn, int bytes) {
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
        (void)hipGetLastError();
}

// Workgroups of kernel `fn` (block threads, lds dynamic bytes) that one CU
// holds at once, registers, LDS and wave slots counted by the runtime (at
// least 1; cached).  A persistent grid larger than this leaves workgroups
// waiting for a slot until others finish, and those run with the device
// part empty.
uint32_t resident_per_cu(const void* fn, uint32_t block, size_t lds);

// Tile shape of the add/sub stream (chosen by the autotuner).
struct AddCfg {
    uint32_t span_digits = 1024;  // digits per wave span (uniform: rounded to whole instances)
    uint32_t tiles_in_flight = 1; // 256-digit tiles whose loads are issued together (1, 2, 4; 8 for ragged adds)
    uint32_t nontemporal = 1;     // streaming loads/stores bypass the caches' LRU
};

// Shape of the convolution multiply (chosen by the autotuner).
struct MulCfg {
    uint32_t waves_per_block = 4; // for integers of <= 512 digits (one team per <= 64 lanes)
    uint32_t rows = 8;            // digits of the row operand per block step (4 or 8)
    uint32_t limb_bits = 32;      // 32: digit convolution, 27: limb convolution (m <= kMul27MaxDigits),
                                  // 8: byte convolution on the matrix cores, opt-in only
                                  // (kMulMfmaMinDigits..kMulMfmaMaxDigits)
};

// sub: |a-b| digit-wise borrow scan; `order` (optional, per instance) swaps
// the operands of the flagged instances.  chunk_states (add only, optional):
// add_chunk_states(cfg, m, n) bytes of scratch; with it, instances longer
// than the span are cut into span-sized chunks, each its own wave, and a
// second pass carries between the chunks.
hipError_t add_uniform(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                       uint64_t n, bool sub, hipStream_t st, uint8_t* order = nullptr,
                       uint8_t* chunk_states = nullptr);
// bytes of chunk states add_uniform needs for (cfg, m, n): 0 when an
// instance fits one span
uint64_t add_chunk_states(const AddCfg& cfg, uint32_t m, uint64_t n);
hipError_t add_csr(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, const int64_t* off,
                   uint64_t n, uint64_t total, bool sub, hipStream_t st);

// Up to kMulSmallDigits digits, batches of at least kMulSmallMinInstances
// instances multiply one lane per instance with the operands in registers
// (fewer instances than that keep the lanes of a team each on part of one
// product, which finishes sooner).
constexpr uint32_t kMulSmallDigits = 64;
constexpr uint64_t kMulSmallMinInstances = 16384;
// above 32 digits (128-row workgroups, two waves per SIMD) from this count:
// enough workgroups to fill the device
constexpr uint64_t kMulSmallMinInstancesWide = 131072;
// Largest m (digits) handled by the one-workgroup-per-team multiply.
constexpr uint32_t kMulMaxDigits = 4096;
// Largest m of the LDS-resident product chain (teams within one wave).
constexpr uint32_t kMulChainMaxDigits = 512;
// Largest m of the register-resident add chain (8 tiles of 256 digits).
constexpr uint32_t kAddChainMaxDigits = 2048;
// rs = (as*bs)^p mod B^m, m <= kMulChainMaxDigits
hipError_t mul_chain_uniform(const MulCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                             uint64_t n, uint32_t p, hipStream_t st, uint32_t* snaps, uint64_t snap_words,
                             uint32_t num_cus);
// rs = p*(as+bs) mod B^m by p chained adds, m <= kAddChainMaxDigits
hipError_t add_chain_uniform(const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m, uint64_t n, uint32_t p,
                             hipStream_t st);
// snaps: per-lane global scratch of mul_snap_words(num_cus) words
hipError_t mul_uniform(const MulCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                       uint64_t n, hipStream_t st, uint32_t* snaps, uint64_t snap_words, uint32_t num_cus);
uint64_t mul_snap_words(uint32_t num_cus);
// Largest m of the 27-bit limb convolution (1024 limbs: no column sum of limb
// products reaches 2^64).
constexpr uint32_t kMul27MaxDigits = 864;
bool mul27_eligible(uint32_t m);
hipError_t mul27_uniform(const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m, uint64_t n, hipStream_t st,
                         uint32_t* snaps, uint64_t snap_words, uint32_t num_cus);
// Byte convolution on the int8 matrix cores (bi_mulmfma.hip), limb_bits 8,
// taken only when a caller sets it (the default multiply and the autotuner
// stay on the multiply-add kernels): m from kMulMfmaMinDigits to
// kMulMfmaMaxDigits (column sums of signed byte products stay below 2^31 far
// beyond that; the bound is the per-wave accumulator tiles, one per 256
// digits).
constexpr uint32_t kMulMfmaMinDigits = 33;
constexpr uint32_t kMulMfmaMaxDigits = 1024;
bool mul_mfma_eligible(uint32_t m);
hipError_t mul_mfma_uniform(const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m, uint64_t n,
                            hipStream_t st, uint32_t num_cus);

// helpers (bi_helpers.hip)
// mode 0: int8 cmp, mode 1: uint8 (a<b), mode 2: uint32 significant digits
hipError_t topscan(int mode, const uint32_t* as, const uint32_t* bs, uint32_t m, uint64_t n, void* out,
                   hipStream_t st);
hipError_t prec_ref(uint32_t* p, uint64_t n, hipStream_t st, bool u64_limbs = false);
hipError_t shift(const uint32_t* as, uint32_t* rs, uint32_t m, uint64_t n, int32_t s, const int32_t* sarr,
                 hipStream_t st);
hipError_t muld(const uint32_t* as, const uint32_t* ds, uint32_t* rs, uint32_t m, uint64_t n, hipStream_t st);
hipError_t quod(const uint32_t* as, const uint32_t* ds, uint32_t* rs, uint32_t* rem, uint32_t m, uint64_t n,
                hipStream_t st);

}  // namespace bi
