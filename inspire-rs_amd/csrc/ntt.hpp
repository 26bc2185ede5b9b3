// LDS-resident negacyclic NTT over one residue of a degree-N polynomial.
//
// Transform definition (bit-identical to the reference's NttContext once the
// reference's Montgomery factor 2^64 is accounted for, src/math/ntt.rs:269-379):
//   forward:  A[k] = sum_j a_j * psi^((2*br(k)+1)*j)  mod q    (k in array order)
//   inverse:  the exact inverse of forward.
// psi is the reference's root (ntt.rs:525-538) and the twiddle table is
// tw[m] = psi^br(m) (ntt.rs:541-583), so slot k of this transform and slot k of
// the reference hold the same evaluation; the reference stores it times
// 2^64 mod q, this code stores it plain.
//
// Structure: Cooley-Tukey (forward) / Gentleman-Sande (inverse) butterflies in
// the reference's stage order, grouped three stages at a time into radix-8
// passes held in registers (8 values per thread, N/8 threads), with one LDS
// round trip between passes.  Butterflies are Harvey's lazy forms with Shoup
// twiddle products; values stay below 4q < 2^32 between stages.
//
// LDS layout: element x lives at x + 4*(x >> 5).  The padding makes every
// pass's access pattern conflict-free (contiguous 32-word runs, or 4-element
// groups that are 16-byte aligned).
#pragma once
#include "modarith.hpp"

namespace inspire {

// Twiddles and constants of one modulus, as the kernels see them.
struct ModDev {
    uint32_t q, q2, qinv;        // q, 2q, -q^-1 mod 2^32
    uint32_t pad_;
    const uint32_t *tw, *twp;    // [N] psi^br(m), Shoup companions
    const uint32_t *itw, *itwp;  // [N] psi^-br(m), Shoup companions
    const uint32_t* twm;         // [N] psi^br(m) 2^32 mod q (Montgomery form, the register transforms')
};

template <int LOGN>
struct NttShape {
    static constexpr int N = 1 << LOGN;
    static constexpr int T = N / 8;            // threads per polynomial
    static constexpr int LDS = N + 4 * (N / 32);  // padded words per polynomial
    static constexpr int R8 = LOGN / 3;        // radix-8 passes
    static constexpr int REM = LOGN % 3;       // leftover stages (0, 1 or 2)
};

__device__ __forceinline__ int lds_ix(int x) { return (int)((unsigned)x + 4 * ((unsigned)x >> 5)); }  // x >= 0

// Twiddle sources: tw(i) -> (psi^br(i), Shoup companion), itw(i) likewise for
// psi^-1.  GlobalTw reads the tables of a ModDev; LdsTw reads (w, w') pairs
// staged in LDS (one 64-bit read per twiddle).
struct GlobalTw {
    const ModDev& M;
    __device__ __forceinline__ uint2 tw(int i) const { return make_uint2(M.tw[i], M.twp[i]); }
    __device__ __forceinline__ uint2 itw(int i) const { return make_uint2(M.itw[i], M.itwp[i]); }
};

struct LdsTw {
    const uint2* fwd;  // [N] (psi^br(i), companion)
    const uint2* inv;  // [N] (psi^-br(i), companion)
    __device__ __forceinline__ uint2 tw(int i) const { return fwd[i]; }
    __device__ __forceinline__ uint2 itw(int i) const { return inv[i]; }
};

// Harvey CT butterfly: X, Y < 4q  ->  X + wY, X - wY  (both < 4q).
__device__ __forceinline__ void ct_bf(uint32_t& X, uint32_t& Y, uint32_t w, uint32_t wp, uint32_t q,
                                      uint32_t q2) {
    uint32_t x = csub(X, q2);
    uint32_t t = shoup_mul(Y, w, wp, q);
    X = x + t;
    Y = x - t + q2;
}

// Harvey GS butterfly: X, Y < 2q  ->  X + Y, w(X - Y)  (both < 2q).
__device__ __forceinline__ void gs_bf(uint32_t& X, uint32_t& Y, uint32_t w, uint32_t wp, uint32_t q,
                                      uint32_t q2) {
    uint32_t s = csub(X + Y, q2);
    uint32_t d = X - Y + q2;
    X = s;
    Y = shoup_mul(d, w, wp, q);
}

// One forward radix-8 pass over stages m0, 2m0, 4m0 (t0 = N / (2 m0)).
template <int LOGN, typename TW>
__device__ __forceinline__ void fwd_pass8(uint32_t* s, int tid, int m0, const ModDev& M, const TW& T) {
    constexpr int N = 1 << LOGN;
    const int t0 = N / (2 * m0), st = t0 >> 2;
    const int i = tid / st, r = tid - i * st;
    const int base = 2 * i * t0 + r;
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = s[lds_ix(base + k * st)];
    {
        const uint2 w = T.tw(m0 + i);
#pragma unroll
        for (int k = 0; k < 4; k++) ct_bf(x[k], x[k + 4], w.x, w.y, M.q, M.q2);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint2 w = T.tw(2 * m0 + 2 * i + h);
        ct_bf(x[4 * h + 0], x[4 * h + 2], w.x, w.y, M.q, M.q2);
        ct_bf(x[4 * h + 1], x[4 * h + 3], w.x, w.y, M.q, M.q2);
    }
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const uint2 w = T.tw(4 * m0 + 4 * i + h);
        ct_bf(x[2 * h], x[2 * h + 1], w.x, w.y, M.q, M.q2);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) s[lds_ix(base + k * st)] = x[k];
}

// Forward leftover pass: 2 stages (radix-4, t = 2, 1) or 1 stage (t = 1).
template <int LOGN, int REM, typename TW>
__device__ __forceinline__ void fwd_tail(uint32_t* s, int tid, const ModDev& M, const TW& T) {
    constexpr int N = 1 << LOGN;
    if constexpr (REM == 2) {
        constexpr int m0 = N / 4;
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int i = tid + g * (N / 8);
            uint4 v = *reinterpret_cast<uint4*>(&s[lds_ix(4 * i)]);
            uint32_t x0 = v.x, x1 = v.y, x2 = v.z, x3 = v.w;
            uint2 w = T.tw(m0 + i);
            ct_bf(x0, x2, w.x, w.y, M.q, M.q2);
            ct_bf(x1, x3, w.x, w.y, M.q, M.q2);
            w = T.tw(2 * m0 + 2 * i);
            ct_bf(x0, x1, w.x, w.y, M.q, M.q2);
            w = T.tw(2 * m0 + 2 * i + 1);
            ct_bf(x2, x3, w.x, w.y, M.q, M.q2);
            *reinterpret_cast<uint4*>(&s[lds_ix(4 * i)]) = make_uint4(x0, x1, x2, x3);
        }
    } else if constexpr (REM == 1) {
        constexpr int m0 = N / 2;
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int i2 = tid + g * (N / 8);  // pair of butterflies 2*i2, 2*i2+1
            uint4 v = *reinterpret_cast<uint4*>(&s[lds_ix(4 * i2)]);
            uint32_t x0 = v.x, x1 = v.y, x2 = v.z, x3 = v.w;
            uint2 w = T.tw(m0 + 2 * i2);
            ct_bf(x0, x1, w.x, w.y, M.q, M.q2);
            w = T.tw(m0 + 2 * i2 + 1);
            ct_bf(x2, x3, w.x, w.y, M.q, M.q2);
            *reinterpret_cast<uint4*>(&s[lds_ix(4 * i2)]) = make_uint4(x0, x1, x2, x3);
        }
    }
}

// In-place forward NTT of the polynomial in LDS `s` by T = N/8 threads
// (tid in [0, T)).  Input values < 4q; output values < 4q (callers reduce).
// `sync` must be a barrier over every thread that shares `s`.
template <int LOGN, typename Sync, typename TW>
__device__ __forceinline__ void ntt_fwd_lds(uint32_t* s, int tid, const ModDev& M, Sync sync, const TW& T) {
    using S = NttShape<LOGN>;
    int m0 = 1;
#pragma unroll
    for (int p = 0; p < S::R8; p++) {
        fwd_pass8<LOGN>(s, tid, m0, M, T);
        sync();
        m0 <<= 3;
    }
    if constexpr (S::REM) {
        fwd_tail<LOGN, S::REM>(s, tid, M, T);
        sync();
    }
}

template <int LOGN, typename Sync>
__device__ __forceinline__ void ntt_fwd_lds(uint32_t* s, int tid, const ModDev& M, Sync sync) {
    ntt_fwd_lds<LOGN>(s, tid, M, sync, GlobalTw{M});
}

// One inverse radix-8 pass over stages t0, 2t0, 4t0.
template <int LOGN, typename TW>
__device__ __forceinline__ void inv_pass8(uint32_t* s, int tid, int t0, const ModDev& M, const TW& T,
                                          bool last, uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    constexpr int N = 1 << LOGN;
    const int i3 = tid / t0, r = tid - i3 * t0;
    const int base = 8 * t0 * i3 + r;
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = s[lds_ix(base + k * t0)];
    const int mA = N / (2 * t0), mB = N / (4 * t0), mC = N / (8 * t0);
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const uint2 w = T.itw(mA + 4 * i3 + h);
        gs_bf(x[2 * h], x[2 * h + 1], w.x, w.y, M.q, M.q2);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint2 w = T.itw(mB + 2 * i3 + h);
        gs_bf(x[4 * h + 0], x[4 * h + 2], w.x, w.y, M.q, M.q2);
        gs_bf(x[4 * h + 1], x[4 * h + 3], w.x, w.y, M.q, M.q2);
    }
    if (!last) {
        const uint2 w = T.itw(mC + i3);
#pragma unroll
        for (int k = 0; k < 4; k++) gs_bf(x[k], x[k + 4], w.x, w.y, M.q, M.q2);
    } else {
        // final stage (m = 1) with the output scale folded in:
        // X' = c (X + Y), Y' = (c psi^-br(1)) (X - Y)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t X = x[k], Y = x[k + 4];
            x[k] = shoup_mul(X + Y, c, cp, M.q);
            x[k + 4] = shoup_mul(X - Y + M.q2, cw, cwp, M.q);
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++) s[lds_ix(base + k * t0)] = x[k];
}

template <int LOGN, int REM, typename TW>
__device__ __forceinline__ void inv_head(uint32_t* s, int tid, const ModDev& M, const TW& T) {
    constexpr int N = 1 << LOGN;
    if constexpr (REM == 2) {  // stages t = 1 (m = N/2) then t = 2 (m = N/4)
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int i = tid + g * (N / 8);
            uint4 v = *reinterpret_cast<uint4*>(&s[lds_ix(4 * i)]);
            uint32_t x0 = v.x, x1 = v.y, x2 = v.z, x3 = v.w;
            uint2 w = T.itw(N / 2 + 2 * i);
            gs_bf(x0, x1, w.x, w.y, M.q, M.q2);
            w = T.itw(N / 2 + 2 * i + 1);
            gs_bf(x2, x3, w.x, w.y, M.q, M.q2);
            w = T.itw(N / 4 + i);
            gs_bf(x0, x2, w.x, w.y, M.q, M.q2);
            gs_bf(x1, x3, w.x, w.y, M.q, M.q2);
            *reinterpret_cast<uint4*>(&s[lds_ix(4 * i)]) = make_uint4(x0, x1, x2, x3);
        }
    } else if constexpr (REM == 1) {  // stage t = 1
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int i2 = tid + g * (N / 8);
            uint4 v = *reinterpret_cast<uint4*>(&s[lds_ix(4 * i2)]);
            uint32_t x0 = v.x, x1 = v.y, x2 = v.z, x3 = v.w;
            uint2 w = T.itw(N / 2 + 2 * i2);
            gs_bf(x0, x1, w.x, w.y, M.q, M.q2);
            w = T.itw(N / 2 + 2 * i2 + 1);
            gs_bf(x2, x3, w.x, w.y, M.q, M.q2);
            *reinterpret_cast<uint4*>(&s[lds_ix(4 * i2)]) = make_uint4(x0, x1, x2, x3);
        }
    }
}

// In-place inverse NTT of the polynomial in LDS.  Input values < 2q.  The
// result is multiplied by the constant c: pass c = n^-1 (with c*psi^-br(1) as
// cw) for the plain inverse, or n^-1 times any extra factor to fold a scaling
// into the last stage.  Output values < 2q.
template <int LOGN, typename Sync, typename TW>
__device__ __forceinline__ void ntt_inv_lds(uint32_t* s, int tid, const ModDev& M, Sync sync, const TW& T,
                                            uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    using S = NttShape<LOGN>;
    static_assert(S::R8 >= 1, "N >= 8");
    if constexpr (S::REM) {
        inv_head<LOGN, S::REM>(s, tid, M, T);
        sync();
    }
    int t0 = 1 << S::REM;
#pragma unroll
    for (int p = 0; p < S::R8; p++) {
        inv_pass8<LOGN>(s, tid, t0, M, T, p == S::R8 - 1, c, cp, cw, cwp);
        sync();
        t0 <<= 3;
    }
}

template <int LOGN, typename Sync>
__device__ __forceinline__ void ntt_inv_lds(uint32_t* s, int tid, const ModDev& M, Sync sync,
                                            uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    ntt_inv_lds<LOGN>(s, tid, M, sync, GlobalTw{M}, c, cp, cw, cwp);
}

}  // namespace inspire
