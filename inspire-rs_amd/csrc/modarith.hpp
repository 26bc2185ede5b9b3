// Exact 32-bit modular arithmetic for moduli q < 2^30 (gfx950 device + host).
//
// Representation: residues are plain u32 values.  Lazily reduced values live
// in [0, 2q) or [0, 4q); every value that leaves a kernel is canonical [0, q).
//  * shoup_mul(x, w, w')  x*w mod q in [0, 2q) for any x < 2^32, with the
//    precomputed w' = floor(w * 2^32 / q) (Harvey / Shoup).
//  * mont_reduce(t)       t * 2^-32 mod q in [0, 2q) for t < q * 2^32.
//    Operands kept in "Montgomery form" (x * 2^32 mod q) turn a product back
//    into a plain value: mont_reduce(a * (b * 2^32)) = a * b.
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#define IR_HD __host__ __device__ __forceinline__
#else
#define IR_HD inline
#endif

namespace inspire {

IR_HD uint32_t umulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

IR_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

// [0, 2m) -> [0, m) for any bound m: unsigned wrap makes x - m huge when x < m.
IR_HD uint32_t csub(uint32_t x, uint32_t m) { return umin32(x, x - m); }

IR_HD uint32_t shoup_mul(uint32_t x, uint32_t w, uint32_t wp, uint32_t q) {
    uint32_t hi = umulhi32(x, wp);
    return x * w - hi * q;
}

IR_HD uint32_t shoup_pre(uint32_t w, uint32_t q) { return (uint32_t)(((uint64_t)w << 32) / q); }

// v mod q for any v < 2^32, with onep = shoup_pre(1, q).
IR_HD uint32_t reduce32(uint32_t v, uint32_t q, uint32_t onep) { return csub(shoup_mul(v, 1u, onep, q), q); }

IR_HD uint32_t mont_reduce(uint64_t t, uint32_t q, uint32_t qinv_neg) {
    uint32_t m = (uint32_t)t * qinv_neg;
    return (uint32_t)((t + (uint64_t)m * q) >> 32);
}

// Host-side helpers used to build constants.
inline uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m) {
    return (uint64_t)(((unsigned __int128)a * b) % m);
}
inline uint64_t powmod64(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mulmod64(r, b, m);
        b = mulmod64(b, b, m);
        e >>= 1;
    }
    return r;
}
inline uint32_t neg_inv32(uint32_t q) {  // -q^{-1} mod 2^32 (q odd)
    uint32_t y = 1;
    for (int i = 0; i < 5; i++) y *= 2u - q * y;  // Newton: doubles correct bits
    return 0u - y;
}

}  // namespace inspire
