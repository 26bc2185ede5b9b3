// Register-resident NTT building blocks for kernels that keep a residue's
// polynomial in registers and exchange through LDS only between passes: V =
// 8 values per thread (N/8 threads per residue, radix-8 passes, Lay) or V = 4
// (N/4 threads, radix-4 passes, Lay4: twice the waves for the same work).
//
// Position layout.  The passes leave every thread on V fixed positions.  The
// inverse transform's last pass and the forward transform's first pass touch
// the same positions {t + k N/V} (the "coefficient positions"); the forward
// transform's last pass and the inverse transform's first pass touch the same
// positions (the "slot positions", V t .. V t + V - 1).
// Coefficient-domain work (CRT composition, gadget digits, automorphisms)
// therefore runs in registers between an inverse and a forward transform, and
// NTT-domain work (multiply-accumulate) between a forward and an inverse one.
//
// Twiddles are Montgomery-form words (w 2^32 mod q) of the forward table,
// normally staged in LDS; a butterfly's product is one Montgomery reduction
// (two 64-bit multiply-adds and one 32-bit multiply).  The inverse transform
// needs no table of its own: psi^-br(m) = -psi^br(3 2^s - 1 - m) for m in
// stage s (2^s <= m < 2^(s+1)), because psi^N = -1 and the stage's
// bit-reversed exponents mirror each other about N; the Gentleman-Sande
// butterfly takes the negation by subtracting the other way round.  Barriers
// wait for LDS traffic only, so global loads issued earlier stay in flight
// across them.
#pragma once
#include <hip/hip_runtime.h>

#include "ntt.hpp"

namespace inspire {
// Barrier over the workgroup for LDS data: this wave's LDS operations have
// completed, then s_barrier.  Global loads issued earlier are not waited for.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Ordering point for an LDS exchange among the lanes of one wave: the LDS
// unit executes a wave's instructions in order, so only the compiler has to
// keep the stores before the loads (wavefront-scope fences around a
// scheduling barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Mod {
    uint32_t q, q2, qinv;
};

// Residue handled by thread tid when each residue has T threads.  With T a
// multiple of the wave size every wave works on one residue, and the index is
// made wave-uniform so that per-residue constants are scalar loads and scalar
// registers rather than per-lane ones.
template <int T>
__device__ __forceinline__ int residue_of(int tid) {
    const int k = tid / T;
    if constexpr (T % 64 == 0) return __builtin_amdgcn_readfirstlane(k);
    return k;
}

// Padded LDS index of element base + e*st, from the padded index of base.
// For the radix-8 passes (st a power of two, base = an aligned block of 8*st
// elements plus r < st) the padding of the eight elements differs from that
// of base by a constant: the addresses are one register plus immediates.
__device__ __forceinline__ int lds_step(int base_ix, int e, int st) { return base_ix + e * st + 4 * ((e * st) >> 5); }

// a * b * 2^-32 mod q in [0, 2q), for a * b < q * 2^32.
__device__ __forceinline__ uint32_t mmul(uint32_t a, uint32_t b, const Mod& M) {
    return mont_reduce((uint64_t)a * b, M.q, M.qinv);
}

// Harvey CT butterfly with a Montgomery-form twiddle: X, Y < 4q -> X + wY, X - wY (< 4q).
__device__ __forceinline__ void ct_m(uint32_t& X, uint32_t& Y, uint32_t wm, const Mod& M) {
    const uint32_t x = csub(X, M.q2);
    const uint32_t t = mmul(Y, wm, M);
    X = x + t;
    Y = x - t + M.q2;
}

// Montgomery reduction with a signed quotient: t * 2^-32 mod q for a signed
// 64-bit t, as a two's-complement word.  |result| <= (|t| + 2^31 q) / 2^32,
// so |t| < 2^31 q gives a result in (-q, q).
__device__ __forceinline__ uint32_t smont(int64_t t, const Mod& M) {
    const int32_t m = (int32_t)((uint32_t)t * M.qinv);
    return (uint32_t)((t + (int64_t)m * (int32_t)M.q) >> 32);
}

// a * b * 2^-32 mod q for a two's-complement word a and 0 <= b < q < 2^31.
__device__ __forceinline__ uint32_t smmul(uint32_t a, uint32_t b, const Mod& M) {
    return smont((int64_t)(int32_t)a * (int32_t)b, M);
}

// Signed CT butterfly: X, Y two's-complement words, X + wY and X - wY with no
// reduction.  A stage widens the bound B of its inputs to B + (B (q - 1) +
// 2^31 q) / 2^32; callers check on the host that every stage stays in int32.
__device__ __forceinline__ void ct_s(uint32_t& X, uint32_t& Y, uint32_t wm, const Mod& M) {
    const uint32_t t = smmul(Y, wm, M);
    Y = X - t;
    X = X + t;
}

// Host: bounds of a signed digit path for modulus q (< 2^29): gadget digits
// below digit_max through logn ct_s stages, nd transformed digits; rec_in > 0:
// a recovered digit is the signed Montgomery sum of a value below rec_in times
// a constant below q and the nd transforms times constants below q; the
// multiply-accumulate sums the transformed and recovered digits times operands
// below q, plus terms below `extra` times constants below q.  True when every
// stage stays inside int32 and every reduced sum inside (-2q, 2q).
inline bool signed_digits_ok(uint64_t q, uint64_t digit_max, int logn, int nd, uint64_t rec_in, uint64_t extra) {
    using u128 = unsigned __int128;
    if (4 * q >= (1ull << 31)) return false;
    auto red = [&](u128 t) { return (uint64_t)((t + ((u128)q << 31)) >> 32); };  // bound of |smont(t)|
    uint64_t B = digit_max;
    for (int s = 0; s < logn; s++) {
        B += red((u128)B * (q - 1));
        if (B >= (1ull << 31)) return false;
    }
    uint64_t last = 0;
    if (rec_in) {
        last = red((u128)rec_in * (q - 1) + (u128)nd * B * (q - 1));
        if (last >= 2 * q) return false;
    }
    return red(((u128)nd * B + last + extra) * (q - 1)) < 2 * q;
}

// Host: bounds of a signed forward transform of inputs below in_max with one
// reduction to (-q, q) after `first` stages (ntt_digits' RED): true when all
// logn stages stay inside int32 and the outputs times a constant below q fit
// one signed Montgomery product (below 2^31 q).
inline bool signed_full_ok(uint64_t q, uint64_t in_max, int logn, int first) {
    using u128 = unsigned __int128;
    if (2 * q >= (1ull << 31)) return false;
    auto red = [&](u128 t) { return (uint64_t)((t + ((u128)q << 31)) >> 32); };
    uint64_t B = in_max;
    for (int s = 0; s < logn; s++) {
        if (s == first) B = red((u128)B * (q - 1));
        B += red((u128)B * (q - 1));
        if (B >= (1ull << 31)) return false;
    }
    return true;
}

template <bool SGN>
__device__ __forceinline__ void ct_b(uint32_t& X, uint32_t& Y, uint32_t wm, const Mod& M) {
    if constexpr (SGN) ct_s(X, Y, wm, M);
    else ct_m(X, Y, wm, M);
}

// Harvey GS butterfly with the mirrored forward twiddle wm standing for -wm:
// X, Y < 2q -> X + Y, (-w)(X - Y) = w(Y - X) (< 2q).
__device__ __forceinline__ void gs_m(uint32_t& X, uint32_t& Y, uint32_t wm, const Mod& M) {
    const uint32_t s = csub(X + Y, M.q2);
    const uint32_t d = Y - X + M.q2;
    X = s;
    Y = mmul(d, wm, M);
}

// Twiddle of the inverse transform's entry m = mS + o of the stage with base
// mS (a power of two): the forward twiddle at the mirrored index 2 mS - 1 - o.
__device__ __forceinline__ uint32_t inv_tw(const uint32_t* tw, int mS, int o) { return tw[2 * mS - 1 - o]; }

template <int LOGN>
struct Lay {
    static constexpr int N = 1 << LOGN, T = N / 8, LP = NttShape<LOGN>::LDS;
    static constexpr int R8 = LOGN / 3, REM = LOGN % 3;
    // first word of the two 4-element groups of a thread's slot positions:
    // eight consecutive slots, so the blocks the last forward passes and the
    // first inverse passes exchange stay inside one wave
    static __device__ __forceinline__ int group(int t, int g) { return 8 * t + 4 * g; }
    // Exchange between forward radix-8 pass p and the next read: the block a
    // pass-(p+1) thread reads was written by N / 8^(p+1) threads of pass p.
    // With at most 64 of them (one wave) the exchange needs no s_barrier.
    static constexpr bool fwd_wave_local(int p) { return (N >> (3 * (p + 1))) <= 64; }
    // Exchange into inverse radix-8 pass p: its blocks of 8 t0 elements, t0 =
    // 2^REM 8^p, were written by the same t0 threads.
    static constexpr bool inv_wave_local(int p) { return ((1 << REM) << (3 * p)) <= 64; }
    // The forward transform passes at least one s_barrier (callers that need
    // one between earlier LDS reads and later writes rely on it otherwise).
    static constexpr bool fwd_has_barrier = !fwd_wave_local(0) && R8 >= 2;
};

template <int LOGN>
__device__ __forceinline__ void load_slots(const uint32_t* p, int t, uint32_t (&v)[8]) {
    using L = Lay<LOGN>;
#pragma unroll
    for (int g = 0; g < 2; g++) {
        const uint4 a = *reinterpret_cast<const uint4*>(p + L::group(t, g));
        v[4 * g + 0] = a.x;
        v[4 * g + 1] = a.y;
        v[4 * g + 2] = a.z;
        v[4 * g + 3] = a.w;
    }
}

// Inverse NTT of one residue, scaled by (c, cw) in the last stage: x at the
// slot positions of thread t -> x at its coefficient positions (< 2q).  rb is
// the residue's [LP] LDS buffer; no thread may still be reading it.
template <int LOGN>
__device__ __forceinline__ void intt1(uint32_t (&x)[8], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                      uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    using L = Lay<LOGN>;
    constexpr int N = L::N;
    if constexpr (L::REM != 0) {
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int i = 2 * t + g;  // the 4-slot group at Lay::group(t, g)
            uint32_t& x0 = x[4 * g + 0];
            uint32_t& x1 = x[4 * g + 1];
            uint32_t& x2 = x[4 * g + 2];
            uint32_t& x3 = x[4 * g + 3];
            gs_m(x0, x1, inv_tw(tw, N / 2, 2 * i), M);
            gs_m(x2, x3, inv_tw(tw, N / 2, 2 * i + 1), M);
            if constexpr (L::REM == 2) {
                const uint32_t w = inv_tw(tw, N / 4, i);
                gs_m(x0, x2, w, M);
                gs_m(x1, x3, w, M);
            }
            *reinterpret_cast<uint4*>(&rb[lds_ix(4 * i)]) = make_uint4(x0, x1, x2, x3);
        }
        wave_lds_sync();  // pass 0's blocks of 8 2^REM elements belong to 2^REM threads
    }
    int t0 = 1 << L::REM;
#pragma unroll
    for (int p = 0; p < L::R8; p++) {
        // thread indices are non-negative: unsigned division is a shift
        const int i3 = (int)((unsigned)t / (unsigned)t0), r = (int)((unsigned)t % (unsigned)t0);
        const int bix = lds_ix(8 * t0 * i3 + r);
        if (!(L::REM == 0 && p == 0)) {
#pragma unroll
            for (int e = 0; e < 8; e++) x[e] = rb[lds_step(bix, e, t0)];
        }
        const int mA = N / (2 * t0), mB = N / (4 * t0), mC = N / (8 * t0);
#pragma unroll
        for (int h = 0; h < 4; h++) gs_m(x[2 * h], x[2 * h + 1], inv_tw(tw, mA, 4 * i3 + h), M);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t w = inv_tw(tw, mB, 2 * i3 + h);
            gs_m(x[4 * h + 0], x[4 * h + 2], w, M);
            gs_m(x[4 * h + 1], x[4 * h + 3], w, M);
        }
        if (p < L::R8 - 1) {
            const uint32_t w = inv_tw(tw, mC, i3);
#pragma unroll
            for (int k = 0; k < 4; k++) gs_m(x[k], x[k + 4], w, M);
#pragma unroll
            for (int e = 0; e < 8; e++) rb[lds_step(bix, e, t0)] = x[e];
            if (L::inv_wave_local(p + 1))
                wave_lds_sync();
            else
                lds_barrier();
        } else {
            // last stage (m = 1) with the output scale: X' = c (X + Y), Y' = cw (X - Y)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t X = x[k], Y = x[k + 4];
                x[k] = shoup_mul(X + Y, c, cp, M.q);
                x[k + 4] = shoup_mul(X - Y + M.q2, cw, cwp, M.q);
            }
        }
        t0 <<= 3;
    }
}

template <int ELL, bool SGN>
__device__ __forceinline__ void fwd_r8(uint32_t (&x)[ELL][8], int m0, int i, const uint32_t* tw, const Mod& M) {
    {
        const uint32_t w = tw[m0 + i];
#pragma unroll
        for (int j = 0; j < ELL; j++)
#pragma unroll
            for (int k = 0; k < 4; k++) ct_b<SGN>(x[j][k], x[j][k + 4], w, M);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint32_t w = tw[2 * m0 + 2 * i + h];
#pragma unroll
        for (int j = 0; j < ELL; j++) {
            ct_b<SGN>(x[j][4 * h + 0], x[j][4 * h + 2], w, M);
            ct_b<SGN>(x[j][4 * h + 1], x[j][4 * h + 3], w, M);
        }
    }
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const uint32_t w = tw[4 * m0 + 4 * i + h];
#pragma unroll
        for (int j = 0; j < ELL; j++) ct_b<SGN>(x[j][2 * h], x[j][2 * h + 1], w, M);
    }
}

// Forward NTT of ELL polys of one residue: x at the coefficient positions of
// thread t -> x at its slot positions (< 4q).  db is [ELL][LP] LDS; no thread
// may still be reading it.  SGN: the inputs and outputs are two's-complement
// words and the butterflies reduce nothing (ct_s): the caller guarantees that
// LOGN stages keep the values inside int32.  RED >= 0 (with SGN): after
// radix-8 pass RED every value is brought back to (-q, q) by a signed
// Montgomery product with r1 = 2^32 mod q, for inputs of full size
// (signed_full_ok bounds the stages on either side).
template <int LOGN, int ELL, bool SGN = false, int RED = -1>
__device__ __forceinline__ void ntt_digits(uint32_t (&x)[ELL][8], int t, uint32_t* db, const uint32_t* tw,
                                           const Mod& M, uint32_t r1 = 0) {
    using L = Lay<LOGN>;
    constexpr int N = L::N, LP = L::LP;
    int m0 = 1;
#pragma unroll
    for (int p = 0; p < L::R8; p++) {
        const int t0 = N / (2 * m0), st = t0 >> 2;
        const int i = (int)((unsigned)t / (unsigned)st), r = (int)((unsigned)t % (unsigned)st);
        const int bix = lds_ix(2 * i * t0 + r);
        if (p > 0) {
#pragma unroll
            for (int j = 0; j < ELL; j++)
#pragma unroll
                for (int e = 0; e < 8; e++) x[j][e] = db[j * LP + lds_step(bix, e, st)];
        }
        fwd_r8<ELL, SGN>(x, m0, i, tw, M);
        if constexpr (SGN && RED >= 0) {
            if (p == RED) {
#pragma unroll
                for (int j = 0; j < ELL; j++)
#pragma unroll
                    for (int e = 0; e < 8; e++) x[j][e] = smmul(x[j][e], r1, M);
            }
        }
        if (!(L::REM == 0 && p == L::R8 - 1)) {
#pragma unroll
            for (int j = 0; j < ELL; j++)
#pragma unroll
                for (int e = 0; e < 8; e++) db[j * LP + lds_step(bix, e, st)] = x[j][e];
            if (L::fwd_wave_local(p))
                wave_lds_sync();
            else
                lds_barrier();
        }
        m0 <<= 3;
    }
    if constexpr (L::REM != 0) {
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int i = 2 * t + g;  // the 4-slot group at Lay::group(t, g)
#pragma unroll
            for (int j = 0; j < ELL; j++) {
                const uint4 v = *reinterpret_cast<const uint4*>(&db[j * LP + lds_ix(4 * i)]);
                x[j][4 * g + 0] = v.x;
                x[j][4 * g + 1] = v.y;
                x[j][4 * g + 2] = v.z;
                x[j][4 * g + 3] = v.w;
            }
            if constexpr (L::REM == 2) {
                const uint32_t w = tw[N / 4 + i], w0 = tw[N / 2 + 2 * i], w1 = tw[N / 2 + 2 * i + 1];
#pragma unroll
                for (int j = 0; j < ELL; j++) {
                    ct_b<SGN>(x[j][4 * g + 0], x[j][4 * g + 2], w, M);
                    ct_b<SGN>(x[j][4 * g + 1], x[j][4 * g + 3], w, M);
                    ct_b<SGN>(x[j][4 * g + 0], x[j][4 * g + 1], w0, M);
                    ct_b<SGN>(x[j][4 * g + 2], x[j][4 * g + 3], w1, M);
                }
            } else {
                const uint32_t w0 = tw[N / 2 + 2 * i], w1 = tw[N / 2 + 2 * i + 1];
#pragma unroll
                for (int j = 0; j < ELL; j++) {
                    ct_b<SGN>(x[j][4 * g + 0], x[j][4 * g + 1], w0, M);
                    ct_b<SGN>(x[j][4 * g + 2], x[j][4 * g + 3], w1, M);
                }
            }
        }
    }
}
// V values per thread (8: the radix-8 layout above, 4: the radix-4 one
// below): the slot positions of thread t are V t .. V t + V - 1.
template <int V>
__device__ __forceinline__ void load_slots_v(const uint32_t* p, int t, uint32_t (&v)[V]) {
#pragma unroll
    for (int g = 0; g < V / 4; g++) {
        const uint4 a = *reinterpret_cast<const uint4*>(p + V * t + 4 * g);
        v[4 * g + 0] = a.x;
        v[4 * g + 1] = a.y;
        v[4 * g + 2] = a.z;
        v[4 * g + 3] = a.w;
    }
}

// The same, non-temporal: data read once (kept out of the caches' way).
template <int V>
__device__ __forceinline__ void load_slots_v_nt(const uint32_t* p, int t, uint32_t (&v)[V]) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int g = 0; g < V / 4; g++) {
        const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p + V * t + 4 * g));
        v[4 * g + 0] = a.x;
        v[4 * g + 1] = a.y;
        v[4 * g + 2] = a.z;
        v[4 * g + 3] = a.w;
    }
}

template <int LOGN>
struct Lay4;
template <int LOGN, int ELL, bool SGN>
__device__ __forceinline__ void ntt4_digits(uint32_t (&x)[ELL][4], int t, uint32_t* db, const uint32_t* tw,
                                            const Mod& M);
template <int LOGN>
__device__ __forceinline__ void intt4(uint32_t (&x)[4], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                      uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp);

template <int LOGN, int V, int ELL, bool SGN = false>
__device__ __forceinline__ void ntt_v(uint32_t (&x)[ELL][V], int t, uint32_t* db, const uint32_t* tw, const Mod& M) {
    if constexpr (V == 8) ntt_digits<LOGN, ELL, SGN>(x, t, db, tw, M);
    else ntt4_digits<LOGN, ELL, SGN>(x, t, db, tw, M);
}

template <int LOGN, int V>
__device__ __forceinline__ void intt_v(uint32_t (&x)[V], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                       uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    if constexpr (V == 8) intt1<LOGN>(x, t, rb, tw, M, c, cp, cw, cwp);
    else intt4<LOGN>(x, t, rb, tw, M, c, cp, cw, cwp);
}

// ---------------------------------------------------------------------------
// Four values per thread (N/4 threads per residue, twice the waves of the
// radix-8 layout): radix-4 passes plus one radix-2 stage when LOGN is odd.
// Coefficient positions {t + e N/4}; slot positions 4t .. 4t + 3.
//
// LDS layouts of the exchanges.  Forward pass p (stride st = N / 4^(p+1))
// has thread t touch logical words 4 st i + r + e st (i = t / st, r = t %
// st); inverse pass p (t0 = 2^REM 4^p) touches 4 t0 i + r + e t0.  With 32
// banks of 4 bytes (lanes of a half wave served together), a stride-8 pass
// puts 4 rows of 32 words under one half wave with 8-word runs at column 8e,
// a stride-2 pass puts the same 4 rows under it with 2-word runs at columns
// 8c + 2e: no padding by rows serves both (the first needs the row offsets
// equal mod 8, the second distinct).  With 64 banks served a whole wave at a
// time, a stride-32 pass puts two rows 128 words apart under one wave, which
// must land 32 banks apart, while the wide passes read 64 consecutive words.
// So each exchange takes a layout its writer and reader are both
// conflict-free in, under either bank model (odd LOGN; checked by
// simulating every access of every pass):
//   kind 1: x + 8 (x >> 5), the exchanges between strides 32 and 8;
//   kind 2: x + 2 (x >> 3), between strides 8 and 2 and around the radix-2
//           stage (read and written there as two 8-byte pairs, which stay
//           contiguous and aligned);
//   kind 3: x + 32 (x >> 7), between strides of 32 and more;
//   kind 0: x + 4 (x >> 5) (lds_ix), for even LOGN, whose stride sets differ.
// In every kind an element's address is the thread's base address plus an
// immediate.
// ---------------------------------------------------------------------------
template <int KIND>
__device__ __forceinline__ int xl_ix(int x) {  // x >= 0
    const unsigned u = (unsigned)x;
    if constexpr (KIND == 1) return (int)(u + 8 * (u >> 5));
    else if constexpr (KIND == 2) return (int)(u + 2 * (u >> 3));
    else if constexpr (KIND == 3) return (int)(u + 32 * (u >> 7));
    else return (int)(u + 4 * (u >> 5));
}
// Index of logical base + e st from the base's index: e st is a multiple of
// the layout's block, or stays inside the base's block.
template <int KIND>
__device__ __forceinline__ int xl_step(int base_ix, int e, int st) {
    if constexpr (KIND == 1) return base_ix + e * st + 8 * ((e * st) >> 5);
    else if constexpr (KIND == 2) return base_ix + e * st + 2 * ((e * st) >> 3);
    else if constexpr (KIND == 3) return base_ix + e * st + 32 * ((e * st) >> 7);
    else return lds_step(base_ix, e, st);
}

template <int LOGN>
struct Lay4 {
    static constexpr int N = 1 << LOGN, T = N / 4;
    static constexpr int R4 = LOGN / 2, REM = LOGN % 2;
    // exchange words per buffer (kinds 1 and 2 hold N + N/4)
    static constexpr int LP = REM ? N + N / 4 : NttShape<LOGN>::LDS;
    // The block a forward pass-(p+1) thread reads was written by the N/4^(p+1)
    // threads of its pass-p block, the block an inverse pass-p thread reads by
    // the same 2^REM 4^p threads: with at most 64 of them (one wave) the
    // exchange needs no s_barrier.
    static constexpr bool fwd_wave_local(int p) { return (N >> (2 * (p + 1))) <= 64; }
    static constexpr bool inv_wave_local(int p) { return ((1 << REM) << (2 * p)) <= 64; }
    // layout of the exchange written by forward pass p, stride st = N / 4^(p+1)
    // (the last one feeds the radix-2 stage)
    static constexpr int fwd_kind(int p) {
        return !REM ? 0 : (N >> (2 * (p + 1))) <= 8 ? 2 : (N >> (2 * (p + 1))) == 32 ? 1 : 3;
    }
    // layout of the exchange written by inverse pass p, t0 = 2 4^p (p = -1: the radix-2 stage)
    static constexpr int inv_kind(int p) { return !REM ? 0 : p <= 0 ? 2 : p == 1 ? 1 : 3; }
};

template <int LOGN, int ELL, bool SGN, int KR, int KW>
__device__ __forceinline__ void ntt4_pass(uint32_t (&x)[ELL][4], int t, int p, int m0, uint32_t* db,
                                          const uint32_t* tw, const Mod& M) {
    using L = Lay4<LOGN>;
    constexpr int N = L::N, LP = L::LP;
    const int t0 = N / (2 * m0), st = t0 >> 1;
    const int i = (int)((unsigned)t / (unsigned)st), r = (int)((unsigned)t % (unsigned)st);
    if (p > 0) {
        const int bix = xl_ix<KR>(2 * i * t0 + r);
#pragma unroll
        for (int j = 0; j < ELL; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) x[j][e] = db[j * LP + xl_step<KR>(bix, e, st)];
    }
    {
        const uint32_t w = tw[m0 + i];
#pragma unroll
        for (int j = 0; j < ELL; j++) {
            ct_b<SGN>(x[j][0], x[j][2], w, M);
            ct_b<SGN>(x[j][1], x[j][3], w, M);
        }
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const uint32_t w = tw[2 * m0 + 2 * i + h];
#pragma unroll
        for (int j = 0; j < ELL; j++) ct_b<SGN>(x[j][2 * h], x[j][2 * h + 1], w, M);
    }
    if (!(L::REM == 0 && p == L::R4 - 1)) {
        const int bix = xl_ix<KW>(2 * i * t0 + r);
#pragma unroll
        for (int j = 0; j < ELL; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) db[j * LP + xl_step<KW>(bix, e, st)] = x[j][e];
        if (L::fwd_wave_local(p))
            wave_lds_sync();
        else
            lds_barrier();
    }
}

template <int LOGN, int ELL, bool SGN, int P>
__device__ __forceinline__ void ntt4_passes(uint32_t (&x)[ELL][4], int t, uint32_t* db, const uint32_t* tw,
                                            const Mod& M) {
    using L = Lay4<LOGN>;
    if constexpr (P < L::R4) {
        constexpr int KR = P > 0 ? L::fwd_kind(P - 1) : 0, KW = L::fwd_kind(P);
        ntt4_pass<LOGN, ELL, SGN, KR, KW>(x, t, P, 1 << (2 * P), db, tw, M);
        ntt4_passes<LOGN, ELL, SGN, P + 1>(x, t, db, tw, M);
    }
}

// Forward NTT of ELL polys of one residue, four values per thread:
// coefficient positions -> slot positions (< 4q; SGN as for ntt_digits).
// db is [ELL][LP] LDS; no thread may still be reading it.
template <int LOGN, int ELL, bool SGN>
__device__ __forceinline__ void ntt4_digits(uint32_t (&x)[ELL][4], int t, uint32_t* db, const uint32_t* tw,
                                            const Mod& M) {
    using L = Lay4<LOGN>;
    constexpr int N = L::N, LP = L::LP;
    ntt4_passes<LOGN, ELL, SGN, 0>(x, t, db, tw, M);
    if constexpr (L::REM != 0) {
        constexpr int KR = L::fwd_kind(L::R4 - 1);
        const uint32_t w0 = tw[N / 2 + 2 * t], w1 = tw[N / 2 + 2 * t + 1];
        const int ix = xl_ix<KR>(4 * t);
#pragma unroll
        for (int j = 0; j < ELL; j++) {
            if constexpr (KR == 2) {  // two 8-byte pairs (the group is 8-byte aligned only)
                const uint2 a = *reinterpret_cast<const uint2*>(&db[j * LP + ix]);
                const uint2 b = *reinterpret_cast<const uint2*>(&db[j * LP + ix + 2]);
                x[j][0] = a.x;
                x[j][1] = a.y;
                x[j][2] = b.x;
                x[j][3] = b.y;
            } else {
                const uint4 v = *reinterpret_cast<const uint4*>(&db[j * LP + ix]);
                x[j][0] = v.x;
                x[j][1] = v.y;
                x[j][2] = v.z;
                x[j][3] = v.w;
            }
            ct_b<SGN>(x[j][0], x[j][1], w0, M);
            ct_b<SGN>(x[j][2], x[j][3], w1, M);
        }
    }
}

template <int LOGN, int P>
__device__ __forceinline__ void intt4_passes(uint32_t (&x)[4], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                             uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    using L = Lay4<LOGN>;
    constexpr int N = L::N;
    if constexpr (P < L::R4) {
        constexpr int KR = L::inv_kind(P - 1), KW = L::inv_kind(P);
        const int t0 = (1 << L::REM) << (2 * P);
        const int i2 = (int)((unsigned)t / (unsigned)t0), r = (int)((unsigned)t % (unsigned)t0);
        if (!(L::REM == 0 && P == 0)) {
            const int bix = xl_ix<KR>(4 * t0 * i2 + r);
#pragma unroll
            for (int e = 0; e < 4; e++) x[e] = rb[xl_step<KR>(bix, e, t0)];
        }
        const int mA = N / (2 * t0), mB = N / (4 * t0);
#pragma unroll
        for (int h = 0; h < 2; h++) gs_m(x[2 * h], x[2 * h + 1], inv_tw(tw, mA, 2 * i2 + h), M);
        if (P < L::R4 - 1) {
            const uint32_t w = inv_tw(tw, mB, i2);
            gs_m(x[0], x[2], w, M);
            gs_m(x[1], x[3], w, M);
            const int bix = xl_ix<KW>(4 * t0 * i2 + r);
#pragma unroll
            for (int e = 0; e < 4; e++) rb[xl_step<KW>(bix, e, t0)] = x[e];
            if (L::inv_wave_local(P + 1))
                wave_lds_sync();
            else
                lds_barrier();
            intt4_passes<LOGN, P + 1>(x, t, rb, tw, M, c, cp, cw, cwp);
        } else {
            // last stage (m = 1) with the output scale: X' = c (X + Y), Y' = cw (X - Y)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint32_t X = x[k], Y = x[k + 2];
                x[k] = shoup_mul(X + Y, c, cp, M.q);
                x[k + 2] = shoup_mul(X - Y + M.q2, cw, cwp, M.q);
            }
        }
    }
}

// Inverse NTT of one residue, four values per thread, scaled by (c, cw) in
// the last stage: slot positions -> coefficient positions (< 2q).  rb is the
// residue's [LP] LDS buffer; no thread may still be reading it.
template <int LOGN>
__device__ __forceinline__ void intt4(uint32_t (&x)[4], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                      uint32_t c, uint32_t cp, uint32_t cw, uint32_t cwp) {
    using L = Lay4<LOGN>;
    constexpr int N = L::N;
    if constexpr (L::REM != 0) {
        constexpr int KW = L::inv_kind(-1);
        gs_m(x[0], x[1], inv_tw(tw, N / 2, 2 * t), M);
        gs_m(x[2], x[3], inv_tw(tw, N / 2, 2 * t + 1), M);
        const int ix = xl_ix<KW>(4 * t);
        if constexpr (KW == 2) {
            *reinterpret_cast<uint2*>(&rb[ix]) = make_uint2(x[0], x[1]);
            *reinterpret_cast<uint2*>(&rb[ix + 2]) = make_uint2(x[2], x[3]);
        } else {
            *reinterpret_cast<uint4*>(&rb[ix]) = make_uint4(x[0], x[1], x[2], x[3]);
        }
        if (L::inv_wave_local(0))
            wave_lds_sync();
        else
            lds_barrier();
    }
    intt4_passes<LOGN, 0>(x, t, rb, tw, M, c, cp, cw, cwp);
}

// ---------------------------------------------------------------------------
// Signed inverse transform (four values per thread) for callers whose inputs
// lie in [0, 2q) and who take outputs in (-q, q), for q < 2^28.  The
// Gentleman-Sande butterflies reduce only the difference:
//   gs_s: X, Y -> X + Y, (-w)(X - Y) = w(Y - X) in (-q, q)  (mirrored twiddle).
// Bounds: the radix-2 stage keeps its sums in [0, 2q) by one conditional
// subtraction; a radix-4 pass takes |x| < 2q to sums below 4q and, in its
// second stage, x[0] (the sum of two sums) below 8q < 2^31 while x[1] stays
// below 2q and the differences below q; x[0] is then reduced to (-q, q) by a
// signed Montgomery product with 2^32 mod q, so every pass starts below 2q.
// The last stage multiplies the sum and the difference (below 8q) by the
// output scales in Montgomery form.  8q < 2^31 also keeps every product
// below 2^31 q, which smont needs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gs_s(uint32_t& X, uint32_t& Y, uint32_t wm, const Mod& M) {
    const uint32_t d = Y - X;
    X = X + Y;
    Y = smmul(d, wm, M);
}

// (-q, q) two's complement -> [0, q)
__device__ __forceinline__ uint32_t sfix(uint32_t x, uint32_t q) {
    return x + ((uint32_t)((int32_t)x >> 31) & q);
}

// Scales of the signed inverse: r1 = 2^32 mod q (reduces x[0] after each pass),
// cm, cwm = the last stage's scales times 2^32 mod q.
struct SInvScale {
    uint32_t r1, cm, cwm;
};

template <int LOGN, int P>
__device__ __forceinline__ void intt4s_passes(uint32_t (&x)[4], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                              const SInvScale& sc) {
    using L = Lay4<LOGN>;
    constexpr int N = L::N;
    if constexpr (P < L::R4) {
        constexpr int KR = L::inv_kind(P - 1), KW = L::inv_kind(P);
        const int t0 = (1 << L::REM) << (2 * P);
        const int i2 = (int)((unsigned)t / (unsigned)t0), r = (int)((unsigned)t % (unsigned)t0);
        if (!(L::REM == 0 && P == 0)) {
            const int bix = xl_ix<KR>(4 * t0 * i2 + r);
#pragma unroll
            for (int e = 0; e < 4; e++) x[e] = rb[xl_step<KR>(bix, e, t0)];
        }
        const int mA = N / (2 * t0), mB = N / (4 * t0);
#pragma unroll
        for (int h = 0; h < 2; h++) gs_s(x[2 * h], x[2 * h + 1], inv_tw(tw, mA, 2 * i2 + h), M);
        if (P < L::R4 - 1) {
            const uint32_t w = inv_tw(tw, mB, i2);
            gs_s(x[0], x[2], w, M);
            gs_s(x[1], x[3], w, M);
            x[0] = smmul(x[0], sc.r1, M);  // below 8q -> (-q, q)
            const int bix = xl_ix<KW>(4 * t0 * i2 + r);
#pragma unroll
            for (int e = 0; e < 4; e++) rb[xl_step<KW>(bix, e, t0)] = x[e];
            if (L::inv_wave_local(P + 1))
                wave_lds_sync();
            else
                lds_barrier();
            intt4s_passes<LOGN, P + 1>(x, t, rb, tw, M, sc);
        } else {
            // last stage (m = 1) with the output scales: X' = c (X + Y), Y' = cw (X - Y)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint32_t X = x[k], Y = x[k + 2];
                x[k] = smmul(X + Y, sc.cm, M);
                x[k + 2] = smmul(X - Y, sc.cwm, M);
            }
        }
    }
}

// Signed inverse NTT of one residue: x in [0, 2q) at the slot positions of
// thread t -> x in (-q, q) at its coefficient positions, scaled as intt4.
template <int LOGN>
__device__ __forceinline__ void intt4_s(uint32_t (&x)[4], int t, uint32_t* rb, const uint32_t* tw, const Mod& M,
                                        const SInvScale& sc) {
    using L = Lay4<LOGN>;
    constexpr int N = L::N;
    if constexpr (L::REM != 0) {
        constexpr int KW = L::inv_kind(-1);
        // radix-2 stage on [0, 2q): the sums kept below 2q, the differences signed
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t& X = x[2 * h];
            uint32_t& Y = x[2 * h + 1];
            const uint32_t d = Y - X;
            X = csub(X + Y, M.q2);
            Y = smmul(d, inv_tw(tw, N / 2, 2 * t + h), M);
        }
        const int ix = xl_ix<KW>(4 * t);
        if constexpr (KW == 2) {
            *reinterpret_cast<uint2*>(&rb[ix]) = make_uint2(x[0], x[1]);
            *reinterpret_cast<uint2*>(&rb[ix + 2]) = make_uint2(x[2], x[3]);
        } else {
            *reinterpret_cast<uint4*>(&rb[ix]) = make_uint4(x[0], x[1], x[2], x[3]);
        }
        if (L::inv_wave_local(0))
            wave_lds_sync();
        else
            lds_barrier();
    }
    intt4s_passes<LOGN, 0>(x, t, rb, tw, M, sc);
}

// Either layout by its number of values per thread.
template <int LOGN, int V>
struct LayV {
    // exchange words per LDS buffer of the transforms (and of any buffer addressed with lds_ix)
    static constexpr int N = 1 << LOGN, T = N / V, LP = V == 8 ? NttShape<LOGN>::LDS : Lay4<LOGN>::LP;
    static __device__ __forceinline__ int group(int t, int g) { return V * t + 4 * g; }
    // the forward transform passes at least one s_barrier
    static constexpr bool fwd_has_barrier = V == 8 ? Lay<LOGN>::fwd_has_barrier : !Lay4<LOGN>::fwd_wave_local(0);
};

}  // namespace inspire
