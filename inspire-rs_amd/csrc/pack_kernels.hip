// InspiRING packing kernels (src/inspiring/inspiring2.rs:926-1041 packing_offline,
// :1185-1222 packing_online), partial packing of gamma column LWEs.
//
// Per query the reference computes, in the NTT domain:
//   step 1   R[i] = tau_{g^i}( gamma^-1 * sum_j NTT(X^(j g^(n-i))) * a_j ),  i < gamma
//   step 2   for i = gamma-2 .. 0:
//              T[i] = G^-1(INTT(R[i+1]));  R[i] += sum_k NTT(T[i][k]) * w_all[i][k]
//   online   b = b_poly + INTT( sum_{i,k} NTT(y_all[i][k]) * NTT(T[i][k]) ),  a = INTT(R[0])
// tau_t in the NTT domain is the reference's apply_automorphism_ntt
// (inspiring2.rs:279-296): residue 0 is permuted, and with two moduli residue 1
// is rebuilt as (residue-0 value in the reference's 2^64 Montgomery form) mod
// q1, read back as a Montgomery value mod q1.  The corrected mode permutes each
// residue instead.
//
// k_pack_chain runs step 2 and the online sum for one query per workgroup:
// the two residues of every transform are processed by the two halves of the
// block at once, the digits live in LDS, R and the online accumulator live in
// registers (8 elements per thread), so T[i] is never written out.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "ntt.hpp"
#include "pack.hpp"

namespace inspire {

struct PackSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// b_poly (respond.rs:350-356): coefficient 0 of column c's b is
// n^-1 * sum_s bhat_c[s].  One block per (query, column, residue).
__global__ void __launch_bounds__(256)
k_pack_b0(const uint32_t* __restrict__ cols, uint32_t* __restrict__ b_poly, uint32_t gamma, CtxDev C,
          uint32_t ninv0, uint32_t ninv1) {
    __shared__ uint64_t sh[4];
    const uint32_t blk = blockIdx.x;
    const uint32_t k = blk % C.K;
    const uint32_t c = (blk / C.K) % gamma;
    const uint32_t qi = blk / (C.K * gamma);
    const size_t KN = (size_t)C.K * C.n;
    const uint32_t* bh = cols + ((size_t)qi * 2 * gamma + 2 * c + 1) * KN + (size_t)k * C.n;
    uint64_t acc = 0;
    for (uint32_t s = threadIdx.x; s < C.n; s += blockDim.x) acc += bh[s];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t q = C.m[k].q;
        uint64_t t = (sh[0] + sh[1] + sh[2] + sh[3]) % q;
        t = (t * (k ? ninv1 : ninv0)) % q;
        b_poly[((size_t)qi * C.K + k) * gamma + c] = (uint32_t)t;
    }
}

// Step 1 before the automorphism:
//   r_pre[q][i][k][s] = sum_j step1[i][j][k][s] * ahat_j[k][s]
// step1 = gamma^-1 NTT(X^(j g^(n-i))) in Montgomery form; kres residues (1
// when the automorphism rebuilds residue 1 from residue 0).
__global__ void __launch_bounds__(256)
k_pack_step1(const uint32_t* __restrict__ cols, const uint32_t* __restrict__ step1, uint32_t* __restrict__ r_pre,
             uint32_t Q, uint32_t gamma, uint32_t kres, CtxDev C) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)Q * kres * C.n) return;
    const uint32_t s = (uint32_t)(gid % C.n);
    const uint32_t k = (uint32_t)((gid / C.n) % kres);
    const uint32_t qi = (uint32_t)(gid / ((size_t)C.n * kres));
    const size_t KN = (size_t)C.K * C.n;
    const uint32_t q = C.m[k].q, qinv = C.m[k].qinv;
    const uint32_t* a = cols + (size_t)qi * 2 * gamma * KN + (size_t)k * C.n + s;
    uint32_t* out = r_pre + (size_t)qi * gamma * KN + (size_t)k * C.n + s;
    for (uint32_t i = 0; i < gamma; i++) {
        const uint32_t* m = step1 + (size_t)i * gamma * KN + (size_t)k * C.n + s;
        uint32_t sum = 0;
        for (uint32_t j0 = 0; j0 < gamma; j0 += 4) {
            uint64_t acc = 0;
            const uint32_t je = j0 + 4 < gamma ? j0 + 4 : gamma;
            for (uint32_t j = j0; j < je; j++) acc += (uint64_t)m[(size_t)j * KN] * a[(size_t)2 * j * KN];
            sum = csub(sum + csub(mont_reduce(acc, q, qinv), q), q);
        }
        out[(size_t)i * KN] = sum;
    }
}

template <int LOGN, int KK>
__global__ void __launch_bounds__(KK * NttShape<LOGN>::T)
k_pack_chain(PackArgs A, CtxDev C, ScaleDev inv) {
    constexpr int N = 1 << LOGN, T = NttShape<LOGN>::T, TT = KK * T, LP = NttShape<LOGN>::LDS;
    extern __shared__ __align__(16) uint32_t lds[];
    const int tid = threadIdx.x;
    const int half = tid / T, ht = tid - half * T;
    const uint32_t qi = blockIdx.x, gamma = A.gamma, ell = C.ell;
    const size_t KN = (size_t)KK * N;
    const uint32_t* rpre = A.r_pre + (size_t)qi * gamma * KN;
    const uint32_t* Y = A.y + A.y_off[qi];
    const uint32_t ymode = A.y_mode[qi], ycount = A.y_count[qi];
    auto buf = [&](uint32_t i, int k) { return lds + ((size_t)i * KK + k) * LP; };
    const ModDev& Mh = C.m[half];
    const uint32_t q0 = C.m[0].q;

    // R_init(i) at (k, s): the reference's apply_automorphism_ntt of r_pre[i].
    auto r_init = [&](uint32_t i, int k, int s) -> uint32_t {
        const uint32_t src = A.perm[(size_t)i * N + s];
        const uint32_t v0 = rpre[(size_t)i * KN + src];
        if (k == 0) return v0;
        if (A.residue_wise) return rpre[(size_t)i * KN + (size_t)N + src];
        const uint32_t q1 = C.m[1].q;
        const uint32_t ref = csub(shoup_mul(v0, A.q0r64, A.q0r64p, q0), q0);
        const uint32_t v1 = csub(shoup_mul(ref, 1u, A.one_q1p, q1), q1);
        return csub(shoup_mul(v1, A.q1r64i, A.q1r64ip, q1), q1);
    };

    uint32_t cur[8], sumb[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        cur[m] = r_init(gamma - 1, k, s);
        sumb[m] = 0;
    }

    for (int i = (int)gamma - 2; i >= 0; i--) {
        // T[i] = G^-1(INTT(R[i+1]))
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int e = tid + m * TT, k = e / N, s = e - k * N;
            buf(0, k)[lds_ix(s)] = cur[m];
        }
        __syncthreads();
        ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, inv.c[half], inv.cp[half], inv.cw[half],
                          inv.cwp[half]);
        uint32_t dig[8 / KK][kMaxEll];
#pragma unroll
        for (int m = 0; m < 8 / KK; m++) {
            const int j = tid + m * TT;
            const uint32_t c0 = csub(buf(0, 0)[lds_ix(j)], q0);
            uint64_t x = c0;
            if (KK == 2) {
                const uint32_t q1 = C.m[1].q;
                const uint32_t c1 = csub(buf(0, 1)[lds_ix(j)], q1);
                const uint32_t a0m = csub(shoup_mul(c0, 1u, C.one_q1p, q1), q1);
                const uint32_t t = csub(shoup_mul(c1 + q1 - a0m, C.q0_inv_q1, C.q0_inv_q1p, q1), q1);
                x = (uint64_t)c0 + (uint64_t)q0 * t;
            }
            for (uint32_t d = 0; d < ell; d++) {  // digit_d = (x / base^d) mod base
                const uint32_t sh = d * C.gadget_log;
                dig[m][d] = sh < 64 ? (uint32_t)((x >> sh) & C.gadget_mask) : 0u;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 8 / KK; m++) {
            const int j = tid + m * TT;
            for (uint32_t d = 0; d < ell; d++)
#pragma unroll
                for (int k = 0; k < KK; k++) {
                    const uint32_t qk = C.m[k].q;
                    const uint32_t v = dig[m][d];
                    buf(d, k)[lds_ix(j)] = v < qk ? v : v % qk;
                }
        }
        __syncthreads();
        for (uint32_t d = 0; d < ell; d++) ntt_fwd_lds<LOGN>(buf(d, half), ht, Mh, PackSync{});
        // R[i] = R_init(i) + sum_d T^[d] * w_all[i][d];  online: += T^[d] * y^[i][d]
        const uint32_t* W = A.w_all + (size_t)i * ell * KN;
        const bool use_y = (uint32_t)i < ycount;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int e = tid + m * TT, k = e / N, s = e - k * N;
            const uint32_t q = C.m[k].q, q2 = 2 * q, qinv = C.m[k].qinv;
            const uint32_t ys = ymode == 0 ? A.perm[(size_t)i * N + s] : (uint32_t)s;
            uint64_t ar = 0, ab = 0;
            for (uint32_t d = 0; d < ell; d++) {
                const uint32_t t = csub(csub(buf(d, k)[lds_ix(s)], q2), q);
                ar += (uint64_t)t * W[((size_t)d * KK + k) * N + s];
                if (use_y) {
                    const size_t yo = ymode == 0 ? ((size_t)d * KK + k) * N + ys
                                                 : (((size_t)i * ell + d) * KK + k) * N + ys;
                    ab += (uint64_t)t * Y[yo];
                }
            }
            cur[m] = csub(r_init((uint32_t)i, k, s) + csub(mont_reduce(ar, q, qinv), q), q);
            if (use_y) sumb[m] = csub(sumb[m] + csub(mont_reduce(ab, q, qinv), q), q);
        }
        __syncthreads();
    }

    // a_hat = INTT(R[0]);  b = b_poly + INTT(sum)
    uint32_t* out = A.out + (size_t)qi * 2 * KN;
    const uint32_t* bp = A.b_poly + (size_t)qi * KK * gamma;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        buf(0, k)[lds_ix(s)] = cur[m];
    }
    __syncthreads();
    ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        out[(size_t)k * N + s] = csub(buf(0, k)[lds_ix(s)], C.m[k].q);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        buf(0, k)[lds_ix(s)] = sumb[m];
    }
    __syncthreads();
    ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        const uint32_t q = C.m[k].q;
        uint32_t v = csub(buf(0, k)[lds_ix(s)], q);
        if ((uint32_t)s < gamma) v = csub(v + bp[(size_t)k * gamma + s], q);
        out[KN + (size_t)k * N + s] = v;
    }
}

static inline unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_pack_b0(const uint32_t* cols, uint32_t* b_poly, uint32_t Q, uint32_t gamma, const CtxDev& C,
                          hipStream_t st) {
    if (!Q) return hipSuccess;
    uint32_t ninv[kMaxK] = {0, 0};
    for (uint32_t k = 0; k < C.K; k++) ninv[k] = (uint32_t)powmod64(C.n, C.m[k].q - 2, C.m[k].q);
    hipLaunchKernelGGL(k_pack_b0, dim3(Q * gamma * C.K), dim3(256), 0, st, cols, b_poly, gamma, C, ninv[0], ninv[1]);
    return hipGetLastError();
}

hipError_t launch_pack_step1(const uint32_t* cols, const uint32_t* step1, uint32_t* r_pre, uint32_t Q,
                             uint32_t gamma, uint32_t kres, const CtxDev& C, hipStream_t st) {
    const size_t total = (size_t)Q * kres * C.n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_pack_step1, dim3(nblk(total, 256)), dim3(256), 0, st, cols, step1, r_pre, Q, gamma, kres, C);
    return hipGetLastError();
}

template <int LOGN, int KK>
static hipError_t chain_t(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv, hipStream_t st) {
    const size_t shmem = (size_t)C.ell * KK * NttShape<LOGN>::LDS * 4;
    auto kern = k_pack_chain<LOGN, KK>;
    if (shmem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(Q), dim3(KK * NttShape<LOGN>::T), shmem, st, A, C, inv);
    return hipGetLastError();
}

template <int LOGN>
static hipError_t chain_k(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv, hipStream_t st) {
    return C.K == 2 ? chain_t<LOGN, 2>(A, Q, C, inv, st) : chain_t<LOGN, 1>(A, Q, C, inv, st);
}

hipError_t launch_pack_chain_scaled(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                    hipStream_t st) {
    if (!Q) return hipSuccess;
    switch (C.logn) {
        case 8: return chain_k<8>(A, Q, C, inv, st);
        case 9: return chain_k<9>(A, Q, C, inv, st);
        case 10: return chain_k<10>(A, Q, C, inv, st);
        case 11: return chain_k<11>(A, Q, C, inv, st);
        case 12: return chain_k<12>(A, Q, C, inv, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace inspire
