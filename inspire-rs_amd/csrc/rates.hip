// Issue-rate probes for the roofs the integer kernels are priced against,
// register-resident so that nothing but the arithmetic is timed:
//   - the Montgomery product of the transforms (mmul: one 32x32->64 product,
//     one 32-bit multiply by -q^-1, one multiply-add by q and the shift);
//   - the signed Montgomery product of the signed digit transforms (smmul);
//   - the 64-bit multiply-add (32x32 + 64) the extract kernel's dot products
//     are made of;
//   - and the streaming read rate of HBM (16-byte loads over a buffer far
//     larger than the caches), the practical roof of the HBM-bound kernels
//     beside the published peak.
// Each thread runs CH independent dependency chains of `iters` operations;
// one vector store per thread, taken only when the chains' XOR equals a value
// the host never passes, keeps the results live.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "reg_ntt.hpp"

namespace inspire {

namespace {

constexpr int kRateChains = 8;

__global__ void __launch_bounds__(256) k_rate_mmul(uint32_t* out, uint32_t iters, Mod M, uint32_t w, uint32_t never) {
    uint32_t x[kRateChains];
#pragma unroll
    for (int c = 0; c < kRateChains; c++) x[c] = (threadIdx.x * 977u + c * 131u + blockIdx.x) % M.q;
    for (uint32_t i = 0; i < iters; i++)
#pragma unroll
        for (int c = 0; c < kRateChains; c++) x[c] = mmul(x[c], w, M);  // x < 2q keeps x w < q 2^32
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < kRateChains; c++) acc ^= x[c];
    if (acc == never) out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ void __launch_bounds__(256) k_rate_smmul(uint32_t* out, uint32_t iters, Mod M, uint32_t w, uint32_t never) {
    uint32_t x[kRateChains];
#pragma unroll
    for (int c = 0; c < kRateChains; c++) x[c] = (threadIdx.x * 977u + c * 131u + blockIdx.x) % M.q;
    for (uint32_t i = 0; i < iters; i++)
#pragma unroll
        for (int c = 0; c < kRateChains; c++) x[c] = smmul(x[c], w, M);  // |x| < q stays in (-q, q)
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < kRateChains; c++) acc ^= x[c];
    if (acc == never) out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ void __launch_bounds__(256) k_rate_mad64(uint32_t* out, uint32_t iters, uint32_t w, uint32_t never) {
    uint64_t x[kRateChains];
#pragma unroll
    for (int c = 0; c < kRateChains; c++) x[c] = threadIdx.x * 977u + c * 131u + blockIdx.x;
    for (uint32_t i = 0; i < iters; i++)
#pragma unroll
        for (int c = 0; c < kRateChains; c++) x[c] = (uint64_t)(uint32_t)x[c] * w + (x[c] >> 32);
    uint32_t acc = 0;
#pragma unroll
    for (int c = 0; c < kRateChains; c++) acc ^= (uint32_t)x[c] ^ (uint32_t)(x[c] >> 32);
    if (acc == never) out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Every 16-byte word of p once, grid-stride, XOR-reduced per thread.
__global__ void __launch_bounds__(256) k_rate_read(const uint4* __restrict__ p, size_t n16, uint32_t* out,
                                                   uint32_t never) {
    uint32_t acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x) {
        const uint4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (acc == never) out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

// Streaming read of `bytes` (a multiple of 16) at src; out: blocks * 256 words.
hipError_t launch_read_probe(const void* src, size_t bytes, uint32_t* out, uint32_t blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_rate_read, dim3(blocks), dim3(256), 0, st, reinterpret_cast<const uint4*>(src), bytes / 16,
                       out, 0xfffffffdu);
    return hipGetLastError();
}

uint64_t rate_ops_per_launch(uint32_t blocks, uint32_t iters) {
    return (uint64_t)blocks * 256 * iters * kRateChains;
}

// kind 0: mmul, 1: smmul, 2: 64-bit multiply-add.  out: blocks * 256 words.
hipError_t launch_rate_probe(int kind, uint32_t* out, uint32_t blocks, uint32_t iters, const ModDev& m,
                             hipStream_t st) {
    const Mod M{m.q, 2 * m.q, m.qinv};
    const uint32_t w = (uint32_t)(((uint64_t)m.q * 3) / 7) | 1u, never = 0xfffffffdu;
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_rate_mmul, dim3(blocks), dim3(256), 0, st, out, iters, M, w, never); break;
        case 1: hipLaunchKernelGGL(k_rate_smmul, dim3(blocks), dim3(256), 0, st, out, iters, M, w, never); break;
        case 2: hipLaunchKernelGGL(k_rate_mad64, dim3(blocks), dim3(256), 0, st, out, iters, w, never); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace inspire
