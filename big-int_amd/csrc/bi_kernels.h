// bi_kernels.h -- launchers of the gfx950 kernels (internal to libbigint).
#pragma once
#include "bi_device.h"

namespace bi {

// Tile shape of the add/sub stream (chosen by the autotuner).
struct AddCfg {
    uint32_t span_digits = 1024;  // digits per wave span (uniform: rounded to whole instances)
    uint32_t tiles_in_flight = 2; // 256-digit tiles whose loads are issued together (1, 2, 4)
    uint32_t nontemporal = 1;     // streaming loads/stores bypass the caches' LRU
};

// Shape of the convolution multiply (chosen by the autotuner).
struct MulCfg {
    uint32_t waves_per_block = 4; // for integers of <= 512 digits (one team per <= 64 lanes)
};

hipError_t add_uniform(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                       uint64_t n, bool sub, hipStream_t st);
hipError_t add_csr(const AddCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, const int64_t* off,
                   uint64_t n, uint64_t total, bool sub, hipStream_t st);

// Largest m (digits) handled by the one-workgroup-per-team multiply.
constexpr uint32_t kMulMaxDigits = 4096;
hipError_t mul_uniform(const MulCfg& cfg, const uint32_t* as, const uint32_t* bs, uint32_t* rs, uint32_t m,
                       uint64_t n, hipStream_t st);

}  // namespace bi
