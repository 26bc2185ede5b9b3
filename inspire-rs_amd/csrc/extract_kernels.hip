// Client-side extract (src/pir/extract.rs) of many responses at once.
//
// One workgroup per response.  Each value of the entry is coefficient c of
// a s + b (RlweCiphertext::decrypt, src/rlwe/enc.rs:90-108): c < ceil(entry /
// 2) of the ciphertext in the packed modes, or coefficient 0 of column
// ciphertext c in the columns mode.  One coefficient of the negacyclic product
// is a dot product, so no transform is needed:
//   (a s)_c = sum_{i <= c} a_i s_{c-i} - sum_{i > c} a_i s_{N+c-i}   (mod q_k)
// The key sits in LDS twice, negated below N: ext[t] = -s_t for t < N and
// s_{t-N} above, so the term of (c, i) is a_i ext[N + c - i] with no sign.
// The waves split residues and halves (quarters with one modulus) of i; a
// pass computes 16 values at once: each lane keeps 16 raw 64-bit sums of its
// products (at most 32 products below 2^58 each), folds them into Montgomery
// residues, and a reduce-scatter across the wave leaves value v on lanes 4v..
// 4v+3.  The waves' partial sums are added in LDS, brought back from the
// 2^-32 the fold left (one product by 2^64 mod q), b's coefficient is added,
// the residues composed (crt_compose_2, src/math/crt.rs:41-50), rounded,
// reduced mod p and written as two bytes per value (reconstruct_entry,
// src/pir/encode_db.rs:186-201).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "../../include/inspire_mi355x.h"
#include "extract.hpp"
#include "modarith.hpp"

namespace inspire {
namespace {

constexpr int kThreads = 256, kWaves = kThreads / 64;
constexpr int V = 16;  // values per pass

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

__device__ __forceinline__ uint32_t addmod(uint32_t a, uint32_t b, uint32_t q) { return csub(a + b, q); }

// A lane's raw sum of products (< 2^63) as acc 2^-32 mod q, canonical:
// acc = hi 2^32 + lo == hi (2^32 mod q) + lo < q 2^32, then one Montgomery step.
__device__ __forceinline__ uint32_t fold(uint64_t acc, uint32_t q, uint32_t qinv, uint32_t r32) {
    const uint64_t t = (acc >> 32) * r32 + (acc & 0xffffffffull);
    return csub(mont_reduce(t, q, qinv), q);
}

// COLS: value c is coefficient 0 of column ciphertext c; else coefficient c
// of the ciphertext.  K: moduli (1 or 2).
// Four workgroups a CU while the doubled key fits 4 in LDS (K d <= 4096 words).
template <int LOGN, int K, bool COLS>
__global__ void __launch_bounds__(kThreads, (K << LOGN) <= 4096 ? 4 : 2) k_extract(ExtractArgs A, CtxDev C) {
    constexpr uint32_t N = 1u << LOGN;
    constexpr uint32_t WPR = kWaves / K;        // waves per residue
    constexpr uint32_t SPAN = N / WPR;          // terms of one value per wave
    constexpr uint32_t TPL = SPAN / 64;         // terms of one value per lane
    constexpr uint32_t SPT = K * N / kThreads;  // key words staged per thread
    static_assert(N >= kThreads && TPL * 64 == SPAN, "shape");
    extern __shared__ uint32_t ext[];  // [K][2N], dynamic: 64 KiB at d = 4096 with two moduli
    __shared__ uint32_t part[kWaves * V];
    __shared__ uint32_t res[V * K];
    const uint32_t qi = blockIdx.x;
    const uint32_t* resp = A.resp + (size_t)qi * A.stride;
    const uint32_t* s = A.sk + (size_t)qi * K * N;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t k = wave / WPR, lo = (wave - k * WPR) * SPAN;
    // every load this response needs first (the key, and a in the packed
    // modes), so their latencies overlap
    uint32_t sv[SPT], av[TPL];
#pragma unroll
    for (uint32_t j = 0; j < SPT; j++) sv[j] = __builtin_nontemporal_load(s + threadIdx.x + j * kThreads);
    if constexpr (!COLS) {
#pragma unroll
        for (uint32_t r = 0; r < TPL; r++) av[r] = __builtin_nontemporal_load(resp + (size_t)k * N + lo + lane + 64 * r);
    }
#pragma unroll
    for (uint32_t j = 0; j < SPT; j++) {
        constexpr uint32_t per = N / kThreads;  // staging steps per residue
        const uint32_t kk = j / per, t = threadIdx.x + (j - kk * per) * kThreads;
        const uint32_t qk = kk ? C.m[1].q : C.m[0].q;
        ext[kk * 2 * N + N + t] = sv[j];
        ext[kk * 2 * N + t] = sv[j] ? qk - sv[j] : 0u;
    }
    __syncthreads();
    const uint32_t q = k ? C.m[1].q : C.m[0].q, qinv = k ? C.m[1].qinv : C.m[0].qinv;
    const uint32_t r32 = k ? C.mont[1] : C.mont[0];
    const uint32_t* ek = ext + k * 2 * N;
    uint8_t* ent = A.entries + (size_t)qi * A.entry_size;

    for (uint32_t c0 = 0; c0 < A.nvals; c0 += V) {
        const uint32_t nc = umin32(V, A.nvals - c0);
        const uint32_t nd = A.ndec > c0 ? umin32(nc, A.ndec - c0) : 0u;  // values decrypted in this pass
        // the b coefficient each value adds, loaded now so the load overlaps the products: b's word
        // k N + coefficient (with b of 2N coefficients per residue these are its first 2N words, as
        // Poly's Add reads them, src/math/poly.rs:710-736)
        uint32_t bw = 0;
        if (threadIdx.x < nd * K) {
            const uint32_t v = threadIdx.x / K, kk = threadIdx.x - v * K, col = c0 + v;
            bw = COLS ? resp[A.col_off + (size_t)col * 2 * K * N + (size_t)K * N + (size_t)kk * N]
                      : resp[A.b_off + (size_t)kk * N + col];
        }
        if (nd) {
            uint32_t x[V];
            if constexpr (COLS) {
                uint64_t acc[V];
#pragma unroll
                for (int v = 0; v < V; v++) acc[v] = 0;
                // column c0 + v for v < nd (the others are not read)
                const uint32_t* a = resp + A.col_off + (size_t)c0 * 2 * K * N + (size_t)k * N + lo + lane;
#pragma unroll 2
                for (uint32_t r = 0; r < TPL; r++) {
                    const uint32_t e = ek[N - (lo + lane + 64 * r)];
                    if (nd == V) {
#pragma unroll
                        for (int v = 0; v < V; v++) acc[v] += (uint64_t)a[(size_t)v * 2 * K * N + 64 * r] * e;
                    } else {
#pragma unroll
                        for (int v = 0; v < V; v++)
                            if ((uint32_t)v < nd) acc[v] += (uint64_t)a[(size_t)v * 2 * K * N + 64 * r] * e;
                    }
                }
#pragma unroll
                for (int v = 0; v < V; v++) x[v] = fold(acc[v], q, qinv, r32);
            } else {
                // coefficients c0 .. c0 + 15, in two halves of 8 (register pressure): ext index
                // N - i + c0 + v <= N + roundup16(nvals) - 1 < 2N, so the values past nd are
                // computed and dropped
#pragma unroll
                for (int half = 0; half < 2; half++) {
                    uint64_t acc[V / 2];
#pragma unroll
                    for (int v = 0; v < V / 2; v++) acc[v] = 0;
#pragma unroll
                    for (uint32_t r = 0; r < TPL; r++) {
                        const uint32_t* e = ek + N - (lo + lane + 64 * r) + c0 + half * (V / 2);
#pragma unroll
                        for (int v = 0; v < V / 2; v++) acc[v] += (uint64_t)av[r] * e[v];
                    }
#pragma unroll
                    for (int v = 0; v < V / 2; v++) x[half * (V / 2) + v] = fold(acc[v], q, qinv, r32);
                }
            }
            // reduce-scatter over the wave: after the step with offset o the
            // lanes with bit o set keep the upper half of their values (the
            // swap is a masked xor, so the halves stay in fixed registers)
#pragma unroll
            for (int h = V / 2, off = 32; h >= 1; h >>= 1, off >>= 1) {
                const uint32_t low = (lane & off) ? 0u : 0xffffffffu;
#pragma unroll
                for (int j = 0; j < h; j++) {
                    const uint32_t d = (x[j] ^ x[j + h]) & low;
                    const uint32_t send = x[j] ^ d, keep = x[j + h] ^ d;
                    x[j] = addmod(keep, __shfl_xor(send, off, 64), q);
                }
            }
            x[0] = addmod(x[0], __shfl_xor(x[0], 2, 64), q);
            x[0] = addmod(x[0], __shfl_xor(x[0], 1, 64), q);
            if ((lane & 3) == 0) part[wave * V + (lane >> 2)] = x[0];  // value lane / 4
        }
        __syncthreads();
        if (threadIdx.x < nd * K) {
            const uint32_t v = threadIdx.x / K, kk = threadIdx.x - v * K;
            const uint32_t qk = kk ? C.m[1].q : C.m[0].q, qkinv = kk ? C.m[1].qinv : C.m[0].qinv;
            uint32_t y = 0;
#pragma unroll
            for (uint32_t w = 0; w < WPR; w++) y = addmod(y, part[(kk * WPR + w) * V + v], qk);
            y = csub(mont_reduce((uint64_t)y * (kk ? A.r64[1] : A.r64[0]), qk, qkinv), qk);  // times 2^32
            res[v * K + kk] = addmod(y, reduce32(bw, qk, kk ? C.onep[1] : C.onep[0]), qk);
        }
        __syncthreads();
        if (threadIdx.x < nc) {
            const uint32_t v = threadIdx.x, col = c0 + v;
            uint32_t val = 0;
            // no column ciphertexts in the columns mode: value 0 stands for every column
            const bool have = col < A.ndec || (!COLS && A.mode == INSPIRE_EXTRACT_COLUMNS);
            if (have) {
                const uint32_t slot = col < A.ndec ? v : 0u;
                uint64_t x = res[slot * K];
                if constexpr (K == 2) {
                    const uint32_t q0 = C.m[0].q, q1 = C.m[1].q;
                    const uint32_t a1 = res[slot * K + 1];
                    const uint32_t a0m = reduce32((uint32_t)x, q1, C.onep[1]);
                    const uint32_t diff = a1 >= a0m ? a1 - a0m : a1 + q1 - a0m;
                    const uint32_t t = csub(shoup_mul(diff, C.q0_inv_q1, C.q0_inv_q1p, q1), q1);
                    x += (uint64_t)q0 * t;
                }
                uint64_t m = (x + A.half_delta) / A.delta;
                m = (A.p & (A.p - 1)) == 0 ? m & (A.p - 1) : m % A.p;
                if (A.mode == INSPIRE_EXTRACT_PACKED) m = mulmod_p(m, A.d_inv, A.p);
                val = (uint32_t)(m & 0xffff);
            }
            const uint32_t o = 2 * col;
            if (o < A.entry_size) ent[o] = (uint8_t)(val & 0xff);
            if (o + 1 < A.entry_size) ent[o + 1] = (uint8_t)(val >> 8);
        }
        __syncthreads();  // the next pass rewrites part and res
    }
}

// Raises a kernel's dynamic LDS limit once (a function-local static is
// initialised exactly once, also under concurrent callers).
template <int LOGN, int K, bool COLS>
hipError_t extract_lds_limit(size_t shmem) {
    static const hipError_t e =
        shmem > 48 * 1024 ? hipFuncSetAttribute((const void*)k_extract<LOGN, K, COLS>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem)
                          : hipSuccess;
    return e;
}

template <int LOGN, int K>
hipError_t launch_k(const ExtractArgs& A, const CtxDev& C, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    constexpr size_t shmem = (size_t)K * 2 * (1u << LOGN) * 4;
    if (A.use_cols) {
        const hipError_t e = extract_lds_limit<LOGN, K, true>(shmem);
        if (e != hipSuccess) return e;
        hipExtLaunchKernelGGL((k_extract<LOGN, K, true>), dim3(A.count), dim3(kThreads), shmem, st, e0, e1, 0, A, C);
    } else {
        const hipError_t e = extract_lds_limit<LOGN, K, false>(shmem);
        if (e != hipSuccess) return e;
        hipExtLaunchKernelGGL((k_extract<LOGN, K, false>), dim3(A.count), dim3(kThreads), shmem, st, e0, e1, 0, A, C);
    }
    return hipGetLastError();
}

template <int LOGN>
hipError_t launch_t(const ExtractArgs& A, const CtxDev& C, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    return C.K == 2 ? launch_k<LOGN, 2>(A, C, st, e0, e1) : launch_k<LOGN, 1>(A, C, st, e0, e1);
}

}  // namespace

hipError_t launch_extract(const ExtractArgs& A, const CtxDev& C, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    if (!A.count || !A.nvals) {
        // an empty batch still completes the profiler's span
        if (e0 && hipEventRecord(e0, st) != hipSuccess) return hipGetLastError();
        if (e1 && hipEventRecord(e1, st) != hipSuccess) return hipGetLastError();
        return hipSuccess;
    }
    switch (C.logn) {
        case 8: return launch_t<8>(A, C, st, e0, e1);
        case 9: return launch_t<9>(A, C, st, e0, e1);
        case 10: return launch_t<10>(A, C, st, e0, e1);
        case 11: return launch_t<11>(A, C, st, e0, e1);
        case 12: return launch_t<12>(A, C, st, e0, e1);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace inspire
