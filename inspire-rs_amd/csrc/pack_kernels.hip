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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "ntt.hpp"
#include "pack.hpp"

namespace inspire {

struct PackSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// b_poly (respond.rs:350-356): coefficient 0 of column c's b is
// n^-1 * sum_s bhat_c[s].  One block per (query, column, residue) writes the
// raw sum (one part); the chain kernel reduces it mod q_k and scales by n^-1.
// bpart: [Q][1][gamma][K] u64.
__global__ void __launch_bounds__(256)
k_pack_b0(const uint32_t* __restrict__ cols, uint64_t* __restrict__ bpart, uint32_t gamma, CtxDev C) {
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
    if (threadIdx.x == 0) bpart[((size_t)qi * gamma + c) * C.K + k] = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---------------------------------------------------------------------------
// Fused front end for gamma = 2^LG columns with gamma^2 | 2n: first dimension,
// step 1 and the b_poly partial sums, one thread per (query, NTT slot s).
// With g = 2n/gamma + 1, g^i = 1 + (2n/gamma) i mod 2n, so the step-1
// monomial exponent j g^(n-i) = j - (2n/gamma) i j and, with w_s the slot's
// evaluation point (w_s^(2n) = 1) and zeta_s = w_s^(2n/gamma),
//   r_pre[i][s] = sum_j (gamma^-1 w_s^j ahat_j[s]) zeta_s^(-ij),
// a gamma-point DFT of the twisted column values, done here in registers.
// zeta_s^-1 = psi^(-(2br(s)+1) 2n/gamma) depends only on c = (2br(s)+1) mod
// gamma.  In the reference's residue handling only residue 0 of r_pre is
// read later, so residue 1 (a-part) is computed in the corrected mode only.
// Column values never leave registers: ahat feeds the DFT, bhat feeds the
// per-wave b sums (bpart [Q][n/64][gamma][K] u64, each part reduced mod q_k).
// No LDS: the b sums are a reduce-scatter across the wave's lanes, which
// leaves room on a CU for this kernel beside another lane's packing chain.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t front_addmod(uint32_t a, uint32_t b, uint32_t q) { return csub(a + b, q); }

// Sums x[v] (v < G <= 64, canonical mod q) over the 64 lanes of the wave;
// returns on lane l the sum of value l / (64 / G).  While more than one value
// remains, each step with offset o keeps the upper half of the values on the
// lanes with bit o set (a masked xor swaps the halves in fixed registers).
template <int G>
__device__ __forceinline__ uint32_t wave_sum_scatter(uint32_t (&x)[G], uint32_t lane, uint32_t q) {
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "G");
    int off = 32;
#pragma unroll
    for (int h = G / 2; h >= 1; h >>= 1, off >>= 1) {
        const uint32_t low = (lane & off) ? 0u : 0xffffffffu;
#pragma unroll
        for (int j = 0; j < h; j++) {
            const uint32_t d = (x[j] ^ x[j + h]) & low;
            const uint32_t send = x[j] ^ d, keep = x[j + h] ^ d;
            x[j] = front_addmod(keep, __shfl_xor(send, off, 64), q);
        }
    }
#pragma unroll
    for (; off >= 1; off >>= 1) x[0] = front_addmod(x[0], __shfl_xor(x[0], off, 64), q);
    return x[0];
}
// Ring degree and residue count of a front-end instantiation: compile-time
// for the common shape (so every load offset is a constant added to a
// wave-uniform base), read from the context otherwise (LOGN = 0, KC = 0).
template <int LOGN, int KC>
struct FrontDims {
    const int n, K;
    __device__ __forceinline__ explicit FrontDims(const CtxDev& C)
        : n(LOGN ? 1 << LOGN : (int)C.n), K(KC ? KC : (int)C.K) {}
};

// One residue of the fused front end for the thread's slot s: the column
// values (a only when WANT_A), the twisted DFT of the a values into r_pre,
// and this wave's b sums.  R, Dq and the tables are wave-uniform bases in
// global memory (buffer views); s is the lane's slot, sb = 4 s its byte offset.
template <int LG, bool WANT_A, int ELL, int LOGN, int KC>
__device__ __forceinline__ void front_residue(int k, const BufU32& R, const BufU32& Dq, const BufU32& rp,
                                              uint64_t* bpart, const BufU32& twist, const uint2* dftw,
                                              const CtxDev& C, uint32_t qi, uint32_t blk, uint32_t sb,
                                              uint32_t cls) {
    constexpr int G = 1 << LG;
    constexpr int CH = G < 8 ? G : 8;  // columns whose digits are loaded together
    const FrontDims<LOGN, KC> F(C);
    const int n = F.n, nb = n / 256;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t KN = (uint32_t)(F.K * n);
    const uint32_t q = C.m[k].q, q2 = 2 * q, qinv = C.m[k].qinv;
    uint32_t ca[ELL], cb[ELL];
#pragma unroll
    for (int i = 0; i < ELL; i++) {
        ca[i] = WANT_A ? R.ld(sb, (uint32_t)((2 * i) * F.K + k) * n) : 0u;
        cb[i] = R.ld(sb, (uint32_t)((2 * i + 1) * F.K + k) * n);
    }
    uint32_t v[G], bv[G];
#pragma unroll
    for (int j0 = 0; j0 < G; j0 += CH) {
        // every digit load of the chunk is issued before the first use
        uint32_t dv[CH][ELL];
#pragma unroll
        for (int jj = 0; jj < CH; jj++)
#pragma unroll
            for (int i = 0; i < ELL; i++) dv[jj][i] = Dq.ldnt(sb, (uint32_t)(((j0 + jj) * ELL + i) * F.K + k) * n);
#pragma unroll
        for (int jj = 0; jj < CH; jj++) {
            const int j = j0 + jj;
            uint64_t aa = 0, bb = 0;
#pragma unroll
            for (int i = 0; i < ELL; i++) {
                if constexpr (WANT_A) aa += (uint64_t)dv[jj][i] * ca[i];
                bb += (uint64_t)dv[jj][i] * cb[i];
            }
            bv[j] = csub(mont_reduce(bb, q, qinv), q);
            if constexpr (WANT_A) {
                const uint32_t a = csub(mont_reduce(aa, q, qinv), q);
                // twist by gamma^-1 w_s^j (Montgomery form) and place at the bit-reversed index
                const uint32_t tw = twist.ld(sb, (uint32_t)(k * G + j) * n);
                int p = 0;
#pragma unroll
                for (int b = 0; b < LG; b++) p |= ((j >> b) & 1) << (LG - 1 - b);
                v[p] = csub(mont_reduce((uint64_t)a * tw, q, qinv), q);
            }
        }
    }
    if constexpr (WANT_A) {
        // radix-2 DIT over the bit-reversed input: X_i = sum_j v_j W^(ij), W = zeta_s^-1
        const uint2* W = dftw + ((size_t)k * (G / 2 > 0 ? G / 2 : 1) + cls) * (G / 2 > 0 ? G / 2 : 1);
#pragma unroll
        for (int len = 2; len <= G; len <<= 1) {
            const int half = len / 2, step = G / len;
#pragma unroll
            for (int b = 0; b < G; b += len)
#pragma unroll
                for (int kk = 0; kk < half; kk++) {
                    uint32_t& X = v[b + kk];
                    uint32_t& Y = v[b + kk + half];
                    if (kk == 0) {
                        const uint32_t x = csub(X, q2), y = csub(Y, q2);
                        X = x + y;
                        Y = x - y + q2;
                    } else {
                        const uint2 w = W[kk * step];
                        ct_bf(X, Y, w.x, w.y, q, q2);
                    }
                }
        }
#pragma unroll
        for (int i = 0; i < G; i++) rp.st(sb, (uint32_t)i * KN + (uint32_t)k * n, csub(csub(v[i], q2), q));
    }
    // this wave's sums of bhat_j over its 64 slots, part blk * 4 + wave
    const uint32_t sum = wave_sum_scatter<G>(bv, lane, q);
    constexpr uint32_t LPV = 64 / G;  // lanes holding each value
    if (lane % LPV == 0) {
        const uint32_t j = lane / LPV, part = blk * 4 + wave;
        bpart[(((size_t)qi * nb * 4 + part) * G + j) * F.K + k] = sum;
    }
}

// InsPIRe^0 in the same launch (a query whose role is 1): the column values
// a_j, b_j of every residue and their sums (the combined ciphertext, summed
// in the NTT domain: the inverse transform is linear mod q_k, so summing
// before it gives the words the reference's sum after it gives) into the
// query's block O [1 + gamma][2][K][N] (combined, then the columns), still
// in the NTT domain; the chain kernel's role-1 workgroups transform them.
template <int LG, int ELL, int LOGN, int KC>
__device__ __forceinline__ void front_plain(const BufU32& R, const BufU32& Dq, uint32_t* O, const CtxDev& C,
                                            uint32_t sb) {
    constexpr int G = 1 << LG;
    constexpr int CH = G < 8 ? G : 8;
    const FrontDims<LOGN, KC> F(C);
    const int n = F.n;
    const uint32_t KN = (uint32_t)(F.K * n);
    const BufU32 Og(O);
    for (int k = 0; k < F.K; k++) {
        const uint32_t q = C.m[k].q, qinv = C.m[k].qinv;
        uint32_t ca[ELL], cb[ELL];
#pragma unroll
        for (int i = 0; i < ELL; i++) {
            ca[i] = R.ld(sb, (uint32_t)((2 * i) * F.K + k) * n);
            cb[i] = R.ld(sb, (uint32_t)((2 * i + 1) * F.K + k) * n);
        }
        uint32_t suma = 0, sumb = 0;
#pragma unroll
        for (int j0 = 0; j0 < G; j0 += CH) {
            uint32_t dv[CH][ELL];
#pragma unroll
            for (int jj = 0; jj < CH; jj++)
#pragma unroll
                for (int i = 0; i < ELL; i++) dv[jj][i] = Dq.ldnt(sb, (uint32_t)(((j0 + jj) * ELL + i) * F.K + k) * n);
#pragma unroll
            for (int jj = 0; jj < CH; jj++) {
                uint64_t aa = 0, bb = 0;
#pragma unroll
                for (int i = 0; i < ELL; i++) {
                    aa += (uint64_t)dv[jj][i] * ca[i];
                    bb += (uint64_t)dv[jj][i] * cb[i];
                }
                const uint32_t a = csub(mont_reduce(aa, q, qinv), q), b = csub(mont_reduce(bb, q, qinv), q);
                const uint32_t oc = (uint32_t)(2 + 2 * (j0 + jj)) * KN + (uint32_t)k * n;
                Og.st(sb, oc, a);
                Og.st(sb, oc + KN, b);
                suma = csub(suma + a, q);
                sumb = csub(sumb + b, q);
            }
        }
        Og.st(sb, (uint32_t)k * n, suma);
        Og.st(sb, KN + (uint32_t)k * n, sumb);
    }
}

// At most 64 VGPRs (eight waves per SIMD): a 256-query launch is 8192 waves,
// which then all fit on the 256 CUs at once instead of leaving a second,
// nearly empty round of workgroups.
template <int LG, bool RW, int ELL, int LOGN, int KC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_pack_front(const uint64_t* __restrict__ qdig, const uint32_t* __restrict__ rows, uint32_t* __restrict__ r_pre,
             uint64_t* __restrict__ bpart, const uint32_t* __restrict__ twist, const uint2* __restrict__ dftw,
             PackRoles roles, uint32_t* __restrict__ out, CtxDev C) {
    constexpr int G = 1 << LG;
    const FrontDims<LOGN, KC> F(C);
    const int n = F.n, nb = n / 256, tid = threadIdx.x;
    const uint32_t qi = blockIdx.x / nb, blk = blockIdx.x % nb;
    const uint32_t s = blk * 256 + tid;
    const size_t KN = (size_t)F.K * n;
    const uint32_t sb = 4 * s;
    const BufU32 R(rows + (size_t)qi * ELL * 2 * KN);
    // the host points a missing shard at a valid one and reports it
    const BufU32 Dq(reinterpret_cast<const uint32_t*>(qdig[qi]));
    if (roles.role && roles.role[qi]) {
        front_plain<LG, ELL, LOGN, KC>(R, Dq, out + roles.qout[qi], C, sb);
        return;
    }
    const uint32_t cls = ((2u * (__brev(s) >> (32 - C.logn)) + 1u) & (G - 1)) >> 1;
    const BufU32 tw(twist), rp(r_pre + (size_t)qi * G * KN);
    front_residue<LG, true, ELL, LOGN, KC>(0, R, Dq, rp, bpart, tw, dftw, C, qi, blk, sb, cls);
    // in the reference's residue handling only residue 0 of r_pre is read later
    for (int k = 1; k < F.K; k++)
        front_residue<LG, RW, ELL, LOGN, KC>(k, R, Dq, rp, bpart, tw, dftw, C, qi, blk, sb, cls);
}

// Step 1 before the automorphism:
//   r_pre[q][i][k][s] = gamma^-1 sum_j NTT(X^(j g^(n-i)))[s] ahat_j[k][s]
// NTT(X^e)[s] = w_s^e (w_s the slot's evaluation point, w_s^(2n) = 1), so with
// u = w_s^(g^(n-i)) the sum is the polynomial sum_j ahat_j u^j, evaluated by
// Horner's rule: gamma - 1 products per (i, s) from one table word u
// (stepu [gamma][K][N], Montgomery form), where a table of every monomial
// would hold gamma^2 words per slot.  kres residues (1 when the automorphism
// rebuilds residue 1 from residue 0).
__global__ void __launch_bounds__(256)
k_pack_step1(const uint32_t* __restrict__ cols, const uint32_t* __restrict__ stepu, uint32_t* __restrict__ r_pre,
             uint32_t Q, uint32_t gamma, uint32_t kres, uint2 ginvm, CtxDev C) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)Q * kres * C.n) return;
    const uint32_t s = (uint32_t)(gid % C.n);
    const uint32_t k = (uint32_t)((gid / C.n) % kres);
    const uint32_t qi = (uint32_t)(gid / ((size_t)C.n * kres));
    const size_t KN = (size_t)C.K * C.n;
    const uint32_t q = C.m[k].q, qinv = C.m[k].qinv;
    const uint32_t gm = k ? ginvm.y : ginvm.x;  // gamma^-1 2^32 mod q_k
    const uint32_t* a = cols + (size_t)qi * 2 * gamma * KN + (size_t)k * C.n + s;
    uint32_t* out = r_pre + (size_t)qi * gamma * KN + (size_t)k * C.n + s;
    for (uint32_t i = 0; i < gamma; i++) {
        const uint32_t u = stepu[(size_t)i * KN + (size_t)k * C.n + s];
        // acc < 3q < 2^32 throughout, so acc * u < q 2^32 for the reduction
        uint32_t acc = a[(size_t)2 * (gamma - 1) * KN];
        for (uint32_t j = gamma - 1; j-- > 0;) acc = mont_reduce((uint64_t)acc * u, q, qinv) + a[(size_t)2 * j * KN];
        out[(size_t)i * KN] = csub(mont_reduce((uint64_t)acc * gm, q, qinv), q);
    }
}

// LDS budget of the chain kernel: twiddle pairs (when they fit), the ell*K
// work polys, the online accumulator and b_poly.
template <int LOGN, int KK>
struct ChainShape {
    static constexpr int N = 1 << LOGN, LP = NttShape<LOGN>::LDS;
    static constexpr bool kLdsTw = (size_t)KK * N * 16 + (size_t)kMaxEll * KK * LP * 4 + (size_t)KK * N * 4 <= 150 * 1024;
    static size_t bytes(uint32_t ell, uint32_t gamma) {
        return (kLdsTw ? (size_t)KK * N * 16 : 0) + (size_t)ell * KK * LP * 4 + (size_t)KK * N * 4 + (size_t)KK * gamma * 4 + 16;
    }
};

template <int LOGN, int KK>
__global__ void __launch_bounds__(KK * NttShape<LOGN>::T, 1)
k_pack_chain(PackArgs A, CtxDev C, ScaleDev inv) {
    using CS = ChainShape<LOGN, KK>;
    constexpr int N = 1 << LOGN, T = NttShape<LOGN>::T, TT = KK * T, LP = NttShape<LOGN>::LDS;
    extern __shared__ __align__(16) uint32_t lds[];
    const int tid = threadIdx.x;
    const int half = tid / T, ht = tid - half * T;
    const uint32_t qi = blockIdx.x, gamma = A.gamma, ell = C.ell;
    const size_t KN = (size_t)KK * N;
    uint2* twf = reinterpret_cast<uint2*>(lds);                    // [KK][N] when kLdsTw
    uint2* twi = twf + (CS::kLdsTw ? KK * N : 0);
    uint32_t* work = reinterpret_cast<uint32_t*>(twi + (CS::kLdsTw ? KK * N : 0));  // [ell][KK][LP]
    uint32_t* sumb = work + (size_t)ell * KK * LP;                 // [KK][N] online accumulator
    uint32_t* bp = sumb + KN;                                       // [KK][gamma] b_poly
    auto buf = [&](uint32_t d, int k) { return work + ((size_t)d * KK + k) * LP; };
    const uint32_t* rpre = A.r_pre + (size_t)qi * gamma * KN;
    const uint32_t* Y = A.y + (A.y_off ? A.y_off[qi] : (uint64_t)qi * ell * KN);
    const uint32_t ymode = A.y_mode ? A.y_mode[qi] : 0u;
    const uint32_t ycount = A.y_count ? A.y_count[qi] : gamma - 1;
    const ModDev& Mh = C.m[half];
    const uint32_t q0 = C.m[0].q;

    if constexpr (CS::kLdsTw) {
        for (int e = tid; e < KK * N; e += TT) {
            const int k = e / N, i = e - k * N;
            twf[e] = make_uint2(C.m[k].tw[i], C.m[k].twp[i]);
            twi[e] = make_uint2(C.m[k].itw[i], C.m[k].itwp[i]);
        }
    }
    // b_poly[k][c] = n^-1 * (sum of the column's bhat over all slots) mod q_k
    for (uint32_t e = tid; e < KK * gamma; e += TT) {
        const uint32_t k = e / gamma, c = e - k * gamma;
        const uint64_t q = C.m[k].q;
        uint64_t s = 0;
        for (uint32_t pb = 0; pb < A.nparts; pb++) s += A.bpart[(((size_t)qi * A.nparts + pb) * gamma + c) * KK + k] % q;
        bp[e] = (uint32_t)(((s % q) * (k ? A.ninv1 : A.ninv0)) % q);
    }
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
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        buf(0, k)[lds_ix(s)] = r_init(gamma - 1, k, s);
        sumb[e] = 0;
    }
    __syncthreads();
    const LdsTw Tl{twf + half * N, twi + half * N};
    const GlobalTw Tg{Mh};

    for (int i = (int)gamma - 2; i >= 0; i--) {
        // T[i] = G^-1(INTT(R[i+1])); R[i+1] sits in buf(0, .)
        if constexpr (CS::kLdsTw)
            ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, Tl, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
        else
            ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, Tg, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
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
#pragma unroll
            for (int d = 0; d < kMaxEll; d++) {  // digit_d = (x / base^d) mod base
                const uint32_t sh = d * C.gadget_log;
                dig[m][d] = ((uint32_t)d < ell && sh < 64) ? (uint32_t)((x >> sh) & C.gadget_mask) : 0u;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 8 / KK; m++) {
            const int j = tid + m * TT;
            for (uint32_t d = 0; d < ell; d++)
#pragma unroll
                for (int k = 0; k < KK; k++) {
                    const uint32_t v = dig[m][d];
                    buf(d, k)[lds_ix(j)] = C.small_digits ? v : reduce32(v, C.m[k].q, C.onep[k]);
                }
        }
        __syncthreads();
        for (uint32_t d = 0; d < ell; d++) {
            if constexpr (CS::kLdsTw) ntt_fwd_lds<LOGN>(buf(d, half), ht, Mh, PackSync{}, Tl);
            else ntt_fwd_lds<LOGN>(buf(d, half), ht, Mh, PackSync{}, Tg);
        }
        // R[i] = R_init(i) + sum_d T^[d] * w_all[i][d]  -> buf(0, .);  online: += T^[d] * y^[i][d]
        const uint32_t* W = A.w_all + (size_t)i * ell * KN;
        const bool use_y = (uint32_t)i < ycount;
#pragma unroll 2
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
            buf(0, k)[lds_ix(s)] = csub(r_init((uint32_t)i, k, s) + csub(mont_reduce(ar, q, qinv), q), q);
            if (use_y) sumb[e] = csub(sumb[e] + csub(mont_reduce(ab, q, qinv), q), q);
        }
        __syncthreads();
    }

    // a_hat = INTT(R[0]);  b = b_poly + INTT(sum)
    uint32_t* out = A.out + (size_t)qi * 2 * KN;
    if constexpr (CS::kLdsTw)
        ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, Tl, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
    else
        ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, Tg, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        out[e] = csub(buf(0, k)[lds_ix(s)], C.m[k].q);
    }
    __syncthreads();
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        buf(0, k)[lds_ix(s)] = sumb[e];
    }
    __syncthreads();
    if constexpr (CS::kLdsTw)
        ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, Tl, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
    else
        ntt_inv_lds<LOGN>(buf(0, half), ht, Mh, PackSync{}, Tg, inv.c[half], inv.cp[half], inv.cw[half], inv.cwp[half]);
    for (int m = 0; m < 8; m++) {
        const int e = tid + m * TT, k = e / N, s = e - k * N;
        const uint32_t q = C.m[k].q;
        uint32_t v = csub(buf(0, k)[lds_ix(s)], q);
        if ((uint32_t)s < gamma) v = csub(v + bp[(size_t)k * gamma + s], q);
        out[KN + e] = v;
    }
}

static inline unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_pack_b0(const uint32_t* cols, uint64_t* bpart, uint32_t Q, uint32_t gamma, const CtxDev& C,
                          hipStream_t st) {
    if (!Q) return hipSuccess;
    hipLaunchKernelGGL(k_pack_b0, dim3(Q * gamma * C.K), dim3(256), 0, st, cols, bpart, gamma, C);
    return hipGetLastError();
}

template <int LG, int ELL, int LOGN, int KC>
static hipError_t front_launch(const uint64_t* qdig, const uint32_t* rows, uint32_t* r_pre, uint64_t* bpart,
                               const uint32_t* twist, const uint2* dftw, const PackRoles& roles, uint32_t* out,
                               uint32_t Q, bool rw, const CtxDev& C, hipStream_t st, LaunchEv ev) {
    const dim3 grid(Q * (C.n / 256));
    if (rw)
        hipExtLaunchKernelGGL((k_pack_front<LG, true, ELL, LOGN, KC>), grid, dim3(256), 0, st, ev.start, ev.stop, 0,
                              qdig, rows, r_pre, bpart, twist, dftw, roles, out, C);
    else
        hipExtLaunchKernelGGL((k_pack_front<LG, false, ELL, LOGN, KC>), grid, dim3(256), 0, st, ev.start, ev.stop, 0,
                              qdig, rows, r_pre, bpart, twist, dftw, roles, out, C);
    return hipGetLastError();
}

template <int LG, int ELL>
static hipError_t front_t(const uint64_t* qdig, const uint32_t* rows, uint32_t* r_pre, uint64_t* bpart,
                          const uint32_t* twist, const uint2* dftw, const PackRoles& roles, uint32_t* out, uint32_t Q,
                          bool rw, const CtxDev& C, hipStream_t st, LaunchEv ev) {
    // d = 2048 with two moduli (the parameter set of every InsPIRe^2 workload here): constant strides
    if constexpr (LG == 4 && ELL == 3)
        if (C.logn == 11 && C.K == 2)
            return front_launch<LG, ELL, 11, 2>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
    return front_launch<LG, ELL, 0, 0>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
}

template <int LG>
static hipError_t front_lg(const uint64_t* qdig, const uint32_t* rows, uint32_t* r_pre, uint64_t* bpart,
                           const uint32_t* twist, const uint2* dftw, const PackRoles& roles, uint32_t* out,
                           uint32_t Q, bool rw, const CtxDev& C, hipStream_t st, LaunchEv ev) {
    switch (C.ell) {
        case 1: return front_t<LG, 1>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 2: return front_t<LG, 2>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 3: return front_t<LG, 3>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 4: return front_t<LG, 4>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pack_front(const uint64_t* qdig, const uint32_t* rows, uint32_t* r_pre, uint64_t* bpart,
                             const uint32_t* twist, const uint2* dftw, uint32_t Q, uint32_t gamma, bool rw,
                             const CtxDev& C, hipStream_t st, LaunchEv ev, const PackRoles& roles, uint32_t* out) {
    if (!Q) return hipSuccess;
    if (roles.role && (!roles.qout || !out)) return hipErrorInvalidValue;
    switch (gamma) {
        case 1: return front_lg<0>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 2: return front_lg<1>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 4: return front_lg<2>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 8: return front_lg<3>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 16: return front_lg<4>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 32: return front_lg<5>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        case 64: return front_lg<6>(qdig, rows, r_pre, bpart, twist, dftw, roles, out, Q, rw, C, st, ev);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pack_step1(const uint32_t* cols, const uint32_t* stepu, uint32_t* r_pre, uint32_t Q,
                             uint32_t gamma, uint32_t kres, uint2 ginvm, const CtxDev& C, hipStream_t st) {
    const size_t total = (size_t)Q * kres * C.n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_pack_step1, dim3(nblk(total, 256)), dim3(256), 0, st, cols, stepu, r_pre, Q, gamma, kres,
                       ginvm, C);
    return hipGetLastError();
}

template <int LOGN, int KK>
static hipError_t chain_t(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv, hipStream_t st) {
    const size_t shmem = ChainShape<LOGN, KK>::bytes(C.ell, A.gamma);
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
                                    hipStream_t st, LaunchEv ev) {
    if (!Q) return hipSuccess;
    if (pack_chain_crt_supported(C)) return launch_pack_chain_crt(A, Q, C, inv, st, ev);
    if (A.roles.role || A.roles.qout) return hipErrorInvalidValue;  // roles are the two-modulus kernel's
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
