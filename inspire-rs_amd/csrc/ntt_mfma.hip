// Forward negacyclic NTT at d = 2048 on the matrix cores (int8 MFMA), the
// transform of k_ntt_fwd computed as two small dense products.
//
// With x'_j = a_j psi^j (the negacyclic twist) and w = psi^2, the slot k of
// the transform holds A[k] = X[br(k)], X[m] = sum_j x'_j w^(j m), the cyclic
// DFT of x' (ntt.rs:269-294 evaluates at psi^(2 br(k) + 1)).  Four-step with
// j = j1 + 32 j2 and m = m2 + 64 m1 (j1, m1 < 32; j2, m2 < 64):
//   Y[j1][m2]  = sum_j2 x'[j1 + 32 j2] w64^(j2 m2)        (32 x 64 times 64 x 64)
//   Y'[j1][m2] = Y[j1][m2] w^(j1 m2)
//   X[m2 + 64 m1] = sum_j1 w32^(m1 j1) Y'[j1][m2]          (32 x 32 times 32 x 64)
// with w64 = w^32 and w32 = w^64.  Residues below 2^28 split into four 7-bit
// limbs, so each product is 16 int8 products whose i32 sums are exact (at
// most 4 * 64 * 127^2 < 2^22 for one limb-sum s = a + b); the sums recombine
// as sum_s S_s 2^(7s), reduced once by a Montgomery step with the constants
// 2^(7s) 2^32 mod q (times the output scale in the second product).
//
// One workgroup of 256 threads (4 waves) per four polynomials of one residue.  The
// matrix-core operand layouts of v_mfma_i32_16x16x{64,32}_i8 (checked on the
// device): lane l holds row (A) or column (B) l % 16 and the k-group l / 16;
// the 16 x 16 result holds column l % 16, rows 4 (l / 16) .. + 3.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "ntt_mfma.hpp"

namespace inspire {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kN = 2048, kThreads = 256;
constexpr int kXRow = 80;   // bytes per row of x' limbs in LDS (64 used, 16-byte aligned rows)
constexpr int kYRow = 40;   // bytes per row of the transposed Y' limbs (32 used, 8-byte aligned rows)

__device__ __forceinline__ uint32_t limb(uint32_t v, int a) { return (v >> (7 * a)) & 127u; }

// Sum of the limb-sum accumulators of one result lane, sum_s S_s 2^(7s),
// reduced: with S_s below (pairs a + b = s) * K * 127^2 (S_3 < 2^22 at
// K = 64), the pairs S_2i + 2^7 S_2i+1 stay below 2^29.1, so four 64-bit
// products with c[2i] = 2^(14 i) 2^32 mod q sum to under 2^58.1 < q 2^32,
// the bound of one Montgomery reduction.
__device__ __forceinline__ uint32_t recombine(const v4i (&acc)[7], int r, const uint32_t* c, uint32_t q,
                                              uint32_t qinv) {
    const uint32_t lo = (uint32_t)acc[0][r] + ((uint32_t)acc[1][r] << 7);
    const uint32_t mid = (uint32_t)acc[2][r] + ((uint32_t)acc[3][r] << 7);
    const uint32_t hi = (uint32_t)acc[4][r] + ((uint32_t)acc[5][r] << 7);
    uint64_t t = (uint64_t)lo * c[0] + (uint64_t)mid * c[2];
    t += (uint64_t)hi * c[4] + (uint64_t)(uint32_t)acc[6][r] * c[6];
    return csub(mont_reduce(t, q, qinv), q);
}

// Polynomials per workgroup: the constant operands (w64 columns, w32 rows)
// are loaded once into registers and reused, and each next polynomial's
// input is loaded while the current one is transformed.
constexpr int kPolysPerGroup = 4;

__global__ void __launch_bounds__(kThreads)
k_ntt_fwd_mfma(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, size_t count,
               MfmaNttTables T, CtxDev C, ScaleDev post) {
    __shared__ __align__(16) uint8_t xl[4 * 32 * kXRow];   // x' limbs [a][j1][j2]
    __shared__ __align__(16) uint8_t yl[4 * 64 * kYRow];   // Y' limbs transposed [a][m2][j1]
    __shared__ __align__(16) uint32_t ob[kN];               // output staging by slot
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int k = blockIdx.x % C.K;
    const size_t poly0 = (size_t)(blockIdx.x / C.K) * kPolysPerGroup;
    const uint32_t q = C.m[k].q, qinv = C.m[k].qinv;
    const int col = lane & 15, kg = lane >> 4;
    const int m2 = 16 * w + col;  // this wave's result columns in both products
    const int j1 = tid & 31, j2b = (tid >> 5) * 8;  // this thread's input row and columns

    // constant operands: w64 limbs (column m2, k = j2 in 16 kg .. + 15; the matrix is
    // symmetric) and w32 limbs (rows m1 = 16 tm + col, k = j1 in 8 kg .. + 7)
    v4i bw[4];
    long aw[2][4];
#pragma unroll
    for (int b = 0; b < 4; b++)
        bw[b] = *reinterpret_cast<const v4i*>(T.w64 + (((size_t)k * 4 + b) * 64 + m2) * 64 + 16 * kg);
#pragma unroll
    for (int tm = 0; tm < 2; tm++)
#pragma unroll
        for (int a = 0; a < 4; a++)
            aw[tm][a] = *reinterpret_cast<const long*>(T.w32 + (((size_t)k * 4 + a) * 32 + 16 * tm + col) * 32 + 8 * kg);
    uint32_t c3[7];  // the recombination constants times the output scale
#pragma unroll
    for (int s = 0; s < 7; s++) {
        const uint32_t c = T.c1[k * 8 + s];
        c3[s] = post.enable ? csub(shoup_mul(c, post.c[k], post.cp[k], q), q) : c;
    }
    uint32_t tw1[8];
#pragma unroll
    for (int e = 0; e < 8; e++) tw1[e] = T.tw1[k * kN + j1 + 32 * (j2b + e)];

    uint32_t in[8];
    auto load_in = [&](size_t poly) {
        const uint32_t* p = src + poly * src_stride + (size_t)k * kN;
#pragma unroll
        for (int e = 0; e < 8; e++) in[e] = p[j1 + 32 * (j2b + e)];
    };
    if (poly0 < count) load_in(poly0);
    for (int pi = 0; pi < kPolysPerGroup; pi++) {
        const size_t poly = poly0 + pi;
        if (poly >= count) break;
        // ---- twist and split: row j1, columns j2b .. j2b + 7
        {
            uint32_t v[8];
#pragma unroll
            for (int e = 0; e < 8; e++) v[e] = csub(mont_reduce((uint64_t)in[e] * tw1[e], q, qinv), q);
#pragma unroll
            for (int a = 0; a < 4; a++) {
                uint32_t lo = 0, hi = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    lo |= limb(v[e], a) << (8 * e);
                    hi |= limb(v[e + 4], a) << (8 * e);
                }
                *reinterpret_cast<uint2*>(&xl[(a * 32 + j1) * kXRow + j2b]) = make_uint2(lo, hi);
            }
        }
        if (poly + 1 < count && pi + 1 < kPolysPerGroup) load_in(poly + 1);  // in flight across the products
        __syncthreads();

        // ---- first product: columns m2, rows j1 = 16 tm + 4 kg + r
#pragma unroll
        for (int tm = 0; tm < 2; tm++) {
            v4i acc[7];
#pragma unroll
            for (int s = 0; s < 7; s++) acc[s] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const v4i ax = *reinterpret_cast<const v4i*>(&xl[(a * 32 + 16 * tm + col) * kXRow + 16 * kg]);
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a + b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ax, bw[b], acc[a + b], 0, 0, 0);
            }
            // Y -> Y' = Y w^(j1 m2), split, stored transposed for the second product
            uint32_t yv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int jr = 16 * tm + 4 * kg + r;
                const uint32_t y = recombine(acc, r, T.c1 + k * 8, q, qinv);
                yv[r] = csub(mont_reduce((uint64_t)y * T.tw2[(k * 32 + jr) * 64 + m2], q, qinv), q);
            }
#pragma unroll
            for (int a = 0; a < 4; a++) {
                uint32_t word = 0;
#pragma unroll
                for (int r = 0; r < 4; r++) word |= limb(yv[r], a) << (8 * r);
                *reinterpret_cast<uint32_t*>(&yl[(a * 64 + m2) * kYRow + 16 * tm + 4 * kg]) = word;
            }
        }
        __syncthreads();

        // ---- second product: columns m2, rows m1 = 16 tm + 4 kg + r
        {
            long by[4];  // Y' limbs, column m2, k = j1 in 8 kg .. + 7
#pragma unroll
            for (int b = 0; b < 4; b++) by[b] = *reinterpret_cast<const long*>(&yl[(b * 64 + m2) * kYRow + 8 * kg]);
#pragma unroll
            for (int tm = 0; tm < 2; tm++) {
                v4i acc[7];
#pragma unroll
                for (int s = 0; s < 7; s++) acc[s] = v4i{0, 0, 0, 0};
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++)
                        acc[a + b] = __builtin_amdgcn_mfma_i32_16x16x32_i8(aw[tm][a], by[b], acc[a + b], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const uint32_t m = (uint32_t)(m2 + 64 * (16 * tm + 4 * kg + r));
                    ob[__brev(m) >> 21] = recombine(acc, r, c3, q, qinv);  // slot br11(m)
                }
            }
        }
        __syncthreads();
        uint32_t* o = dst + poly * dst_stride + (size_t)k * kN;
#pragma unroll
        for (int e = 0; e < kN / kThreads; e++) o[tid + e * kThreads] = ob[tid + e * kThreads];
        // the next polynomial's split writes xl (read before the second barrier) and its
        // products write yl and ob only after its first barrier, which every thread passes
        // after these reads of ob
    }
}

}  // namespace

bool ntt_mfma_supported(const CtxDev& C) {
    if (C.logn != 11 || !C.mfma.tw1) return false;
    for (uint32_t k = 0; k < C.K; k++)
        if (C.m[k].q >= (1u << 28)) return false;  // four 7-bit limbs
    return true;
}

hipError_t launch_ntt_fwd_mfma(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, size_t count,
                               const MfmaNttTables& T, const CtxDev& C, const ScaleDev& sc, hipStream_t st) {
    const size_t blocks = (count + kPolysPerGroup - 1) / kPolysPerGroup * C.K;
    if (!blocks) return hipSuccess;
    if (!ntt_mfma_supported(C) || blocks > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ntt_fwd_mfma, dim3((unsigned)blocks), dim3(kThreads), 0, st, src, src_stride, dst, dst_stride,
                       count, T, C, sc);
    return hipGetLastError();
}

}  // namespace inspire

namespace inspire {

MfmaNttHost ntt_mfma_tables(uint32_t K, const uint64_t* q, const uint64_t* psi) {
    MfmaNttHost H;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    H.o_tw1 = 0;
    H.o_tw2 = al(H.o_tw1 + (size_t)K * kN * 4);
    H.o_w64 = al(H.o_tw2 + (size_t)K * 32 * 64 * 4);
    H.o_w32 = al(H.o_w64 + (size_t)K * 4 * 64 * 64);
    H.o_c1 = al(H.o_w32 + (size_t)K * 4 * 32 * 32);
    H.bytes.assign(al(H.o_c1 + (size_t)K * 8 * 4), 0);
    uint32_t* tw1 = reinterpret_cast<uint32_t*>(H.bytes.data() + H.o_tw1);
    uint32_t* tw2 = reinterpret_cast<uint32_t*>(H.bytes.data() + H.o_tw2);
    int8_t* w64 = reinterpret_cast<int8_t*>(H.bytes.data() + H.o_w64);
    int8_t* w32 = reinterpret_cast<int8_t*>(H.bytes.data() + H.o_w32);
    uint32_t* c1 = reinterpret_cast<uint32_t*>(H.bytes.data() + H.o_c1);
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t m = q[k], r32 = (1ull << 32) % m, w = mulmod64(psi[k], psi[k], m);
        for (int j = 0; j < kN; j++) tw1[k * kN + j] = (uint32_t)mulmod64(powmod64(psi[k], j, m), r32, m);
        for (int j1 = 0; j1 < 32; j1++)
            for (int m2 = 0; m2 < 64; m2++)
                tw2[(k * 32 + j1) * 64 + m2] = (uint32_t)mulmod64(powmod64(w, (uint64_t)j1 * m2, m), r32, m);
        const uint64_t w64r = powmod64(w, 32, m), w32r = powmod64(w, 64, m);
        for (int a = 0; a < 64; a++)
            for (int b = 0; b < 64; b++) {
                const uint64_t v = powmod64(w64r, (uint64_t)a * b, m);
                for (int l = 0; l < 4; l++) w64[(((size_t)k * 4 + l) * 64 + a) * 64 + b] = (int8_t)((v >> (7 * l)) & 127);
            }
        for (int a = 0; a < 32; a++)
            for (int b = 0; b < 32; b++) {
                const uint64_t v = powmod64(w32r, (uint64_t)a * b, m);
                for (int l = 0; l < 4; l++) w32[(((size_t)k * 4 + l) * 32 + a) * 32 + b] = (int8_t)((v >> (7 * l)) & 127);
            }
        for (int s = 0; s < 7; s++) c1[k * 8 + s] = (uint32_t)mulmod64((1ull << (7 * s)) % m, r32, m);
    }
    return H;
}

}  // namespace inspire
