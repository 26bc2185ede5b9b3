// Batched client-side extract (src/pir/extract.rs) on the device.
#pragma once
#include <cstddef>
#include <cstdint>

#include "../../include/inspire_mi355x.h"
#include "kernels.hpp"

namespace inspire {

// Kernel-side description of an extract batch (extract_kernels.hip).
struct ExtractArgs {
    const uint32_t* resp;     // [count][stride] responses
    const uint32_t* sk;       // [count][K][N] secret keys
    uint8_t* entries;         // [count][entry_size]
    uint64_t stride;          // words per response
    uint32_t count;
    uint32_t mode;            // INSPIRE_EXTRACT_*
    uint32_t entry_size;
    uint32_t nvals;           // column values an entry takes: ceil(entry_size / 2)
    uint32_t ndec;            // values with a decryption of their own (the rest are 0, or copies of value 0)
    uint32_t use_cols;        // 1: value c is coefficient 0 of column ciphertext c
    uint32_t b_off;           // words from a response to its ciphertext's b (K N)
    uint32_t col_off;         // words from a response to its first column ciphertext (K N + K b_dim)
    uint64_t delta, half_delta, p;
    uint64_t d_inv;           // extract_packed's multiplier: d^-1 mod p, or 1
    uint32_t r64[kMaxK];      // 2^64 mod q_k
};

// start/stop: events the launch records around the kernel (null: none).
hipError_t launch_extract(const ExtractArgs& A, const CtxDev& C, hipStream_t st, hipEvent_t start = nullptr,
                          hipEvent_t stop = nullptr);

// Runs a device batch (inspire_extract_batch_device).
void extract_batch(inspire_ctx* ctx, const inspire_extract_batch& b);

}  // namespace inspire
