// Client-side extract (src/pir/extract.rs) of many responses at once.
//
// One workgroup per response.  Each value of the entry is coefficient c of
// a s + b (RlweCiphertext::decrypt, src/rlwe/enc.rs:90-108) for c < ceil(entry
// / 2) in the packed modes, or coefficient 0 of column ciphertext c in the
// columns mode.  A single coefficient of the negacyclic product is a dot
// product, so no transform is needed:
//   (a s)_c = sum_{i <= c} a_i s_{c-i} - sum_{i > c} a_i s_{N+c-i}   (mod q_k)
// A wave computes one (value, residue) dot product: lane l takes i = l + 64 r,
// reading a coalesced from HBM and s from LDS (consecutive lanes read
// consecutive words).  The products are Montgomery-reduced (a s 2^-32 in
// [0, 2q)), negated terms enter as 2q - x, and the wave's sum is brought back
// with one multiplication by 2^32 mod q.  Then the residues are composed
// (crt_compose_2, src/math/crt.rs:41-50), rounded, reduced mod p and written
// as two bytes per value (reconstruct_entry, src/pir/encode_db.rs:186-201).
#include <hip/hip_runtime.h>

#include "../../include/inspire_mi355x.h"
#include "extract.hpp"
#include "modarith.hpp"

namespace inspire {
namespace {

constexpr int kThreads = 256, kWaves = kThreads / 64;
constexpr int kChunk = 256;  // values finished per pass over the LDS result slots

// a b mod p for a, b < p (p < 2^58 here: p <= q)
__device__ __forceinline__ uint64_t mulmod_p(uint64_t a, uint64_t b, uint64_t p) {
    if (p <= 0xffffffffull) return a * b % p;
    uint64_t r = 0;
    while (b) {
        if (b & 1) { r += a; if (r >= p) r -= p; }
        a += a;
        if (a >= p) a -= p;
        b >>= 1;
    }
    return r;
}

template <int LOGN>
__global__ void __launch_bounds__(kThreads) k_extract(ExtractArgs A, CtxDev C) {
    constexpr uint32_t N = 1u << LOGN;
    __shared__ uint32_t sk[kMaxK * N];
    __shared__ uint32_t res[kChunk * kMaxK];
    const uint32_t K = C.K;
    const uint32_t qi = blockIdx.x;
    const uint32_t* resp = A.resp + (size_t)qi * A.stride;
    const uint32_t* s = A.sk + (size_t)qi * K * N;
    for (uint32_t i = threadIdx.x; i < K * N; i += kThreads) sk[i] = s[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint8_t* ent = A.entries + (size_t)qi * A.entry_size;

    for (uint32_t c0 = 0; c0 < A.nvals; c0 += kChunk) {
        const uint32_t nc = umin32(kChunk, A.nvals - c0);
        const uint32_t nd = A.ndec > c0 ? umin32(nc, A.ndec - c0) : 0u;
        for (uint32_t j = wave; j < nd * K; j += kWaves) {
            const uint32_t v = j / K, k = j - v * K, col = c0 + v;
            const uint32_t* ct = A.use_cols ? resp + A.col_off + (size_t)col * 2 * K * N : resp;
            const uint32_t* bp = A.use_cols ? ct + K * N : resp + A.b_off;
            const uint32_t coef = A.use_cols || A.mode == INSPIRE_EXTRACT_COLUMNS ? 0u : col;
            const uint32_t q = C.m[k].q, qinv = C.m[k].qinv;
            const uint32_t* a = ct + (size_t)k * N;
            const uint32_t* sr = sk + k * N;
            uint64_t acc = 0;
#pragma unroll 4
            for (uint32_t i = lane; i < N; i += 64) {
                const uint32_t x = mont_reduce((uint64_t)a[i] * sr[(coef - i) & (N - 1)], q, qinv);
                acc += i > coef ? (uint64_t)(2 * q - x) : (uint64_t)x;  // X^N = -1
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (lane == 0) {
                // b's word k N + coef: with a b of 2N coefficients per residue
                // these are the first 2N words, as Poly's Add reads them
                // (src/math/poly.rs:710-736); reduce it for residue k
                const uint64_t as = (acc % q) * C.mont[k] % q;
                res[v * kMaxK + k] = (uint32_t)((as + bp[(size_t)k * N + coef] % q) % q);
            }
        }
        __syncthreads();
        for (uint32_t v = threadIdx.x; v < nc; v += kThreads) {
            const uint32_t col = c0 + v;
            uint32_t val = 0;
            const bool have = col < A.ndec || (!A.use_cols && A.mode == INSPIRE_EXTRACT_COLUMNS);
            if (have) {
                const uint32_t slot = col < A.ndec ? v : 0u;  // no column ciphertexts: value 0 everywhere
                uint64_t x = res[slot * kMaxK];
                if (K == 2) {
                    const uint32_t q0 = C.m[0].q, q1 = C.m[1].q;
                    const uint32_t a1 = res[slot * kMaxK + 1];
                    const uint32_t a0m = (uint32_t)(x % q1);
                    const uint32_t diff = a1 >= a0m ? a1 - a0m : a1 + q1 - a0m;
                    const uint32_t t = (uint32_t)((uint64_t)diff * C.q0_inv_q1 % q1);
                    x += (uint64_t)q0 * t;
                }
                uint64_t m = (x + A.half_delta) / A.delta % A.p;
                if (A.mode == INSPIRE_EXTRACT_PACKED) m = mulmod_p(m, A.d_inv, A.p);
                val = (uint32_t)(m & 0xffff);
            }
            const uint32_t o = 2 * col;
            if (o < A.entry_size) ent[o] = (uint8_t)(val & 0xff);
            if (o + 1 < A.entry_size) ent[o + 1] = (uint8_t)(val >> 8);
        }
        __syncthreads();  // the next chunk rewrites res
    }
}

template <int LOGN>
hipError_t launch_t(const ExtractArgs& A, const CtxDev& C, hipStream_t st) {
    hipLaunchKernelGGL(k_extract<LOGN>, dim3(A.count), dim3(kThreads), 0, st, A, C);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_extract(const ExtractArgs& A, const CtxDev& C, hipStream_t st) {
    if (!A.count || !A.nvals) return hipSuccess;
    switch (C.logn) {
        case 8: return launch_t<8>(A, C, st);
        case 9: return launch_t<9>(A, C, st);
        case 10: return launch_t<10>(A, C, st);
        case 11: return launch_t<11>(A, C, st);
        case 12: return launch_t<12>(A, C, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace inspire
