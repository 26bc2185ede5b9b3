// gfx950 kernels of the InsPIRe respond path.
//
// Layouts (u32 residues, canonical [0, q) at every kernel boundary):
//   poly            [K][N]                  residue-major, like the reference's Poly.coeffs
//   NTT-domain poly [K][N]                  plain values (not Montgomery form), slot k
//                                            = evaluation at psi^(2 br(k) + 1)
//   DB digits       [shard][col][ell][K][N] NTT of the gadget digits of Delta*h
//   RGSW b-rows     [query][ell][2][K][N]   rows ell..2ell-1 in NTT domain, Montgomery
//                                            form (value * 2^32 mod q)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "ntt_mfma.hpp"
#include "ntt.hpp"
#include "reg_ntt.hpp"

namespace inspire {

struct BlockSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// ---------------------------------------------------------------------------
// Batched NTT over [count][K][N] polynomials: one block per (poly, residue).
// Forward: optional post-scale by a per-residue constant (Shoup pair), used to
// produce Montgomery form (c = 2^32 mod q) or the reference's 2^64 form.
// Inverse: the per-residue constants (c, c psi^-br(1)) fold n^-1 and any
// pre-scale into the last stage.
// ---------------------------------------------------------------------------
template <int LOGN>
__global__ void __launch_bounds__(NttShape<LOGN>::T)
k_ntt_fwd(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, CtxDev C, ScaleDev post) {
    using S = NttShape<LOGN>;
    __shared__ __align__(16) uint32_t s[S::LDS];
    const int tid = threadIdx.x;
    const int k = blockIdx.x % C.K;
    const size_t poly = blockIdx.x / C.K;
    const ModDev& M = C.m[k];
    const uint32_t* p = src + poly * src_stride + (size_t)k * S::N;
    uint32_t* o = dst + poly * dst_stride + (size_t)k * S::N;
#pragma unroll
    for (int j = 0; j < 8; j++) s[lds_ix(tid + j * S::T)] = p[tid + j * S::T];
    __syncthreads();
    ntt_fwd_lds<LOGN>(s, tid, M, BlockSync{});
    const uint32_t c = post.c[k], cp = post.cp[k];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        uint32_t x = s[lds_ix(tid + j * S::T)];
        x = post.enable ? shoup_mul(x, c, cp, M.q) : csub(x, M.q2);
        o[tid + j * S::T] = csub(x, M.q);
    }
}

template <int LOGN>
__global__ void __launch_bounds__(NttShape<LOGN>::T)
k_ntt_inv(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, CtxDev C, ScaleDev fin) {
    using S = NttShape<LOGN>;
    __shared__ __align__(16) uint32_t s[S::LDS];
    const int tid = threadIdx.x;
    const int k = blockIdx.x % C.K;
    const size_t poly = blockIdx.x / C.K;
    const ModDev& M = C.m[k];
    const uint32_t* p = src + poly * src_stride + (size_t)k * S::N;
    uint32_t* o = dst + poly * dst_stride + (size_t)k * S::N;
#pragma unroll
    for (int j = 0; j < 8; j++) s[lds_ix(tid + j * S::T)] = p[tid + j * S::T];
    __syncthreads();
    ntt_inv_lds<LOGN>(s, tid, M, BlockSync{}, fin.c[k], fin.cp[k], fin.cw[k], fin.cwp[k]);
#pragma unroll
    for (int j = 0; j < 8; j++) o[tid + j * S::T] = csub(s[lds_ix(tid + j * S::T)], M.q);
}

// ---------------------------------------------------------------------------
// Gadget decomposition of composite coefficients (external_product.rs:16-37).
// x in [0, q) is cut into ell base-2^glog digits, least significant first:
// digit_i = (x / base^i) mod base, so bits above base^ell are dropped exactly
// as the reference's `val % base; val /= base` loop drops them.  A digit d is
// stored as d mod q_k in every residue (Poly::set_coeff, poly.rs:375-388).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t crt_compose(const CtxDev& C, uint32_t a0, uint32_t a1) {
    if (C.K == 1) return a0;
    const uint32_t q1 = C.m[1].q;
    // crt_compose_2 (crt.rs:41-51): a0 + q0 * (((a1 - a0 mod q1) * q0^-1) mod q1)
    uint32_t a0m = csub(shoup_mul(a0, 1u, C.one_q1p, q1), q1);
    uint32_t diff = a1 + q1 - a0m;
    uint32_t t = csub(shoup_mul(diff, C.q0_inv_q1, C.q0_inv_q1p, q1), q1);
    return (uint64_t)a0 + (uint64_t)C.m[0].q * t;
}

__device__ __forceinline__ uint32_t digit_of(const CtxDev& C, uint64_t x, int i, int k) {
    const uint32_t sh = i * C.gadget_log;
    const uint32_t v = sh < 64 ? (uint32_t)((x >> sh) & C.gadget_mask) : 0u;
    return C.small_digits ? v : reduce32(v, C.m[k].q, C.onep[k]);
}

// Database preprocessing: h -> Delta*h (trivial_encrypt, rlwe/enc.rs:174-182)
// -> composite -> gadget digits, coefficient domain.  in: [P][K][N] canonical,
// out: [P][ell][K][N].
__global__ void __launch_bounds__(256)
k_db_digits(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t P, CtxDev C) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= P * C.n) return;
    const size_t p = gid / C.n;
    const uint32_t j = (uint32_t)(gid - p * C.n);
    const uint32_t* h = in + p * C.K * C.n;
    uint32_t b[kMaxK];
    for (uint32_t k = 0; k < C.K; k++)
        b[k] = csub(shoup_mul(h[k * C.n + j], C.delta[k], C.deltap[k], C.m[k].q), C.m[k].q);
    const uint64_t x = crt_compose(C, b[0], C.K > 1 ? b[1] : 0u);
    uint32_t* o = out + p * C.ell * C.K * C.n;
    for (uint32_t i = 0; i < C.ell; i++)
        for (uint32_t k = 0; k < C.K; k++) o[(i * C.K + k) * C.n + j] = digit_of(C, x, (int)i, (int)k);
}

// Decompose both components of RLWE ciphertexts: in [P][2][K][N] ->
// out [P][2][ell][K][N] (a digits, then b digits).
__global__ void __launch_bounds__(256)
k_rlwe_digits(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t P, CtxDev C) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= P * 2 * C.n) return;
    const size_t pc = gid / C.n;  // (ciphertext, component)
    const uint32_t j = (uint32_t)(gid - pc * C.n);
    const uint32_t* h = in + pc * C.K * C.n;
    const uint64_t x = crt_compose(C, h[j], C.K > 1 ? h[C.n + j] : 0u);
    uint32_t* o = out + pc * C.ell * C.K * C.n;
    for (uint32_t i = 0; i < C.ell; i++)
        for (uint32_t k = 0; k < C.K; k++) o[(i * C.K + k) * C.n + j] = digit_of(C, x, (int)i, (int)k);
}

// ---------------------------------------------------------------------------
// First dimension (external product of the trivial encryption (0, Delta*h)
// with RGSW rows ell..2ell-1, external_product.rs:131-147; the a-digits are 0):
//   out[q][c].{a,b}[k][s] = sum_i D[shard][c][i][k][s] * row[q][i].{a,b}[k][s]
// D plain, rows in Montgomery form, so one Montgomery reduction of the 64-bit
// sum yields the plain value.  Each thread owns 4 consecutive slots.
// out: out_stride polys per query; column c of query q goes to polys
// q*out_stride + out_off + 2c (a) and + 1 (b), each K*N words.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_first_dim(const uint64_t* __restrict__ qdig, const uint32_t* __restrict__ rows, uint32_t* __restrict__ out,
            uint32_t Q, uint32_t cols, uint32_t out_stride, uint32_t out_off, CtxDev C) {
    const uint32_t n4 = C.n / 4;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)Q * cols * C.K * n4;
    if (gid >= total) return;
    const uint32_t s4 = (uint32_t)(gid % n4);
    size_t r = gid / n4;
    const uint32_t k = (uint32_t)(r % C.K); r /= C.K;
    const uint32_t c = (uint32_t)(r % cols);
    const uint32_t qi = (uint32_t)(r / cols);
    const uint32_t* dig = reinterpret_cast<const uint32_t*>(qdig[qi]);
    if (!dig) return;  // a missing shard: reported by the host
    const size_t KN = (size_t)C.K * C.n;
    const uint4* D = reinterpret_cast<const uint4*>(dig + (size_t)c * C.ell * KN + k * C.n) + s4;
    const uint4* R = reinterpret_cast<const uint4*>(rows + (size_t)qi * C.ell * 2 * KN + k * C.n) + s4;
    uint64_t aa[4] = {0, 0, 0, 0}, bb[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < C.ell; i++) {
        const uint4 d = D[i * KN / 4];
        const uint4 ra = R[(2 * i) * KN / 4];
        const uint4 rb = R[(2 * i + 1) * KN / 4];
        aa[0] += (uint64_t)d.x * ra.x; aa[1] += (uint64_t)d.y * ra.y;
        aa[2] += (uint64_t)d.z * ra.z; aa[3] += (uint64_t)d.w * ra.w;
        bb[0] += (uint64_t)d.x * rb.x; bb[1] += (uint64_t)d.y * rb.y;
        bb[2] += (uint64_t)d.z * rb.z; bb[3] += (uint64_t)d.w * rb.w;
    }
    const uint32_t q = C.m[k].q, qi_ = C.m[k].qinv;
    uint4 oa, ob;
    oa.x = csub(mont_reduce(aa[0], q, qi_), q); oa.y = csub(mont_reduce(aa[1], q, qi_), q);
    oa.z = csub(mont_reduce(aa[2], q, qi_), q); oa.w = csub(mont_reduce(aa[3], q, qi_), q);
    ob.x = csub(mont_reduce(bb[0], q, qi_), q); ob.y = csub(mont_reduce(bb[1], q, qi_), q);
    ob.z = csub(mont_reduce(bb[2], q, qi_), q); ob.w = csub(mont_reduce(bb[3], q, qi_), q);
    uint32_t* o = out + ((size_t)qi * out_stride + out_off + 2 * c) * KN + k * C.n;
    reinterpret_cast<uint4*>(o)[s4] = oa;
    reinterpret_cast<uint4*>(o + KN)[s4] = ob;
}

// General external product MAC over all 2*ell rows:
//   out[p].{a,b} = sum_r dig[p][r] * rgsw[p][r].{a,b}, r < 2 ell,
// dig [P][2ell][K][N] plain NTT, rgsw [P][2ell][2][K][N] Montgomery NTT.
__global__ void __launch_bounds__(256)
k_ext_mac(const uint32_t* __restrict__ dig, const uint32_t* __restrict__ rgsw, uint32_t* __restrict__ out,
          size_t P, CtxDev C) {
    const uint32_t n4 = C.n / 4;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= P * C.K * n4) return;
    const uint32_t s4 = (uint32_t)(gid % n4);
    const size_t r = gid / n4;
    const uint32_t k = (uint32_t)(r % C.K);
    const size_t p = r / C.K;
    const size_t KN = (size_t)C.K * C.n;
    const uint32_t rows = 2 * C.ell;
    const uint4* D = reinterpret_cast<const uint4*>(dig + p * rows * KN + k * C.n) + s4;
    const uint4* G = reinterpret_cast<const uint4*>(rgsw + p * rows * 2 * KN + k * C.n) + s4;
    uint64_t aa[4] = {0, 0, 0, 0}, bb[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < rows; i++) {
        const uint4 d = D[i * KN / 4];
        const uint4 ga = G[(2 * i) * KN / 4];
        const uint4 gb = G[(2 * i + 1) * KN / 4];
        aa[0] += (uint64_t)d.x * ga.x; aa[1] += (uint64_t)d.y * ga.y;
        aa[2] += (uint64_t)d.z * ga.z; aa[3] += (uint64_t)d.w * ga.w;
        bb[0] += (uint64_t)d.x * gb.x; bb[1] += (uint64_t)d.y * gb.y;
        bb[2] += (uint64_t)d.z * gb.z; bb[3] += (uint64_t)d.w * gb.w;
    }
    const uint32_t q = C.m[k].q, qi_ = C.m[k].qinv;
    uint4 oa, ob;
    oa.x = csub(mont_reduce(aa[0], q, qi_), q); oa.y = csub(mont_reduce(aa[1], q, qi_), q);
    oa.z = csub(mont_reduce(aa[2], q, qi_), q); oa.w = csub(mont_reduce(aa[3], q, qi_), q);
    ob.x = csub(mont_reduce(bb[0], q, qi_), q); ob.y = csub(mont_reduce(bb[1], q, qi_), q);
    ob.z = csub(mont_reduce(bb[2], q, qi_), q); ob.w = csub(mont_reduce(bb[3], q, qi_), q);
    uint32_t* o = out + p * 2 * KN + k * C.n;
    reinterpret_cast<uint4*>(o)[s4] = oa;
    reinterpret_cast<uint4*>(o + KN)[s4] = ob;
}

// InsPIRe^0 fold (respond.rs:102-109): resp[q][0] = sum_c resp[q][1 + c],
// resp = [Q][1 + cols][2][K][N] coefficient domain.
__global__ void __launch_bounds__(256)
k_sum_cols(uint32_t* __restrict__ resp, uint32_t Q, uint32_t cols, CtxDev C) {
    const size_t KN = (size_t)C.K * C.n;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)Q * 2 * KN) return;
    const size_t qi = gid / (2 * KN);
    const size_t e = gid - qi * 2 * KN;
    const uint32_t k = (uint32_t)((e % KN) / C.n);
    const uint32_t q = C.m[k].q;
    uint32_t* base = resp + qi * (1 + cols) * 2 * KN;
    uint32_t acc = 0;
    for (uint32_t c = 0; c < cols; c++) acc = csub(acc + base[(1 + c) * 2 * KN + e], q);
    base[e] = acc;
}

// ---------------------------------------------------------------------------
// Seeded RLWE expansion: Poly::from_seed_moduli (poly.rs:278-281).  ChaCha20Rng
// (rand_chacha 0.3: 64-bit block counter in words 12-13, stream 0 in 14-15)
// yields u64 draws v_t = word[2t] | word[2t+1] << 32; residue k takes the next
// N draws accepted by rand 0.8's UniformInt<u64>::sample_single(0, q_k):
// lo(v * q_k) <= (q_k << lz(q_k)) - 1, value hi(v * q_k).  Residue 1 starts at
// the draw after residue 0's N-th acceptance, so the block scans the stream in
// chunks of 256 ChaCha blocks (2048 draws) with a block-wide prefix sum over
// the accept flags.
// One block of 256 threads per seed; out: [K][N] per seed.
// ---------------------------------------------------------------------------
#define CC_QR(a, b, c, d)                         \
    a += b; d ^= a; d = (d << 16) | (d >> 16);   \
    c += d; b ^= c; b = (b << 12) | (b >> 20);   \
    a += b; d ^= a; d = (d << 8) | (d >> 24);    \
    c += d; b ^= c; b = (b << 7) | (b >> 25);

__device__ __forceinline__ void chacha20_block(const uint32_t key[8], uint64_t ctr, uint32_t out[16]) {
    uint32_t x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                      key[4], key[5], key[6], key[7], (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
    uint32_t s[16];
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = x[i];
#pragma unroll
    for (int r = 0; r < 10; r++) {
        CC_QR(x[0], x[4], x[8], x[12]) CC_QR(x[1], x[5], x[9], x[13])
        CC_QR(x[2], x[6], x[10], x[14]) CC_QR(x[3], x[7], x[11], x[15])
        CC_QR(x[0], x[5], x[10], x[15]) CC_QR(x[1], x[6], x[11], x[12])
        CC_QR(x[2], x[7], x[8], x[13]) CC_QR(x[3], x[4], x[9], x[14])
    }
#pragma unroll
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

// Exclusive block-wide prefix sum of v over 256 threads; returns the total.
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t& excl, uint32_t* sh) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        uint32_t t = sh[w];
        before += (w < wv) ? t : 0u;
        total += t;
    }
    excl = before + x - v;
    __syncthreads();
    return total;
}

constexpr int kExpandMaxChunks = 64;

__global__ void __launch_bounds__(256)
k_expand_seeds(const uint32_t* __restrict__ seeds, uint32_t* __restrict__ out, size_t out_stride,
               uint32_t* __restrict__ err, CtxDev C) {
    __shared__ uint32_t sh[4];
    __shared__ uint32_t s_last;
    const int tid = threadIdx.x;
    uint32_t key[8];
#pragma unroll
    for (int i = 0; i < 8; i++) key[i] = seeds[(size_t)blockIdx.x * 8 + i];
    uint32_t* dst = out + (size_t)blockIdx.x * out_stride;
    uint32_t k = 0, filled = 0;
    uint64_t pos = 0;  // first draw not yet consumed by residue k
    for (int chunk = 0; chunk < kExpandMaxChunks && k < C.K; chunk++) {
        const uint64_t c0 = (uint64_t)chunk * 2048;
        uint32_t w[16];
        chacha20_block(key, (uint64_t)chunk * 256 + tid, w);
        uint64_t v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = (uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32);
        // A chunk can finish residue 0 and continue with residue 1.
        while (k < C.K) {
            const uint64_t q = C.m[k].q;
            const uint64_t zone = (q << __clzll(q)) - 1;
            uint32_t flags = 0, cnt = 0;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint64_t idx = c0 + 8 * tid + j;
                const bool ok = idx >= pos && v[j] * q <= zone;
                flags |= (uint32_t)ok << j;
                cnt += ok;
            }
            uint32_t excl;
            const uint32_t total = block_scan256(cnt, excl, sh);
            const uint32_t need = C.n - filled;
            uint32_t rank = excl;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (flags >> j & 1) {
                    if (rank < need) dst[(size_t)k * C.n + filled + rank] = (uint32_t)__umul64hi(v[j], q);
                    if (rank == need - 1) s_last = 8 * tid + j;
                    rank++;
                }
            }
            __syncthreads();
            if (total >= need) {
                pos = c0 + s_last + 1;
                filled = 0;
                k++;
                __syncthreads();
                continue;  // next residue, same chunk
            }
            filled += total;
            pos = c0 + 2048;
            break;
        }
    }
    if (k < C.K && tid == 0) atomicOr(err, 1u);
}

// ---------------------------------------------------------------------------
// Seeded RGSW rows in one pass (d = 2048, 256 threads = the expansion's draw
// lanes = the NTT's N/8): row i's a is expanded from its seed into LDS as
// above, b is read from bsrc + i * KN straight into registers, and per
// residue a and b are transformed together by the register-resident radix-8
// NTT (eight values per thread, Montgomery-form twiddles from the global
// table), scaled to Montgomery form, and written to rows + i * 2KN (a) and
// + KN (b), each thread's eight consecutive slots as two 16-byte stores.
// LDS: three polynomial buffers (residue 0's a in 0, residue 1's in 2; the
// transform of residue k exchanges through buffers k, k + 1 once its a is in
// registers), 27 KB, which leaves a CU running the packing chain's
// workgroup room for one of these.
// Residues of a: a row whose query reads only residue 0 of a (InspiRING
// packing through the fused front end in the reference's residue handling)
// expands and transforms that residue alone: its draws come first in the
// stream, so the residue-0 words are the same, and residue 1's a is left
// unwritten.  Row r expands KA = (a_full && a_full[r / ell]) ? K : a_res.
// SGN: the transforms run on two's-complement words with butterflies that
// reduce nothing (one reduction of every value after six stages), when the
// host's bound check (signed_full_ok) passes for every modulus.
// ---------------------------------------------------------------------------
template <bool SGN>
__global__ void __launch_bounds__(256)
k_rows_seeded(const uint32_t* __restrict__ seeds, const uint32_t* __restrict__ bsrc, uint32_t* __restrict__ rows,
              uint32_t* __restrict__ err, CtxDev C, ScaleDev sc, const uint32_t* __restrict__ a_full,
              uint32_t a_res, uint32_t ell) {
    using S = NttShape<11>;
    constexpr int N = S::N, LP = S::LDS;
    static_assert(kMaxK == 2, "a buffers at 0 and 2");
    __shared__ __align__(16) uint32_t poly[3][LP];  // a of residue 0, exchange, a of residue 1
    __shared__ uint32_t sh[4];
    __shared__ uint32_t s_last;
    const int tid = threadIdx.x;
    const size_t KN = (size_t)C.K * N;
    uint32_t key[8];
#pragma unroll
    for (int i = 0; i < 8; i++) key[i] = seeds[(size_t)blockIdx.x * 8 + i];
    // Draws are generated in chunks of up to 256 ChaCha blocks (8 draws per
    // thread).  For the last residue the chunk is sized to the draws still
    // needed plus a margin (a short chunk just continues in the next one), so
    // the rejected draws of the last residue do not cost a whole chunk.
    const uint32_t KA = (a_full && a_full[blockIdx.x / ell]) ? C.K : a_res;  // residues of a expanded
    uint32_t k = 0, filled = 0;
    uint64_t pos = 0, c0 = 0;  // first unconsumed draw; first draw of the chunk
    for (int chunk = 0; chunk < kExpandMaxChunks && k < KA; chunk++) {
        uint32_t active = 256;
        if (k == KA - 1) {
            const uint32_t need_now = N - filled;
            active = min(256u, (need_now + need_now / 8 + 64 + 7) / 8);
        }
        const bool on = (uint32_t)tid < active;
        uint32_t w[16];
        if (on) chacha20_block(key, c0 / 8 + tid, w);
        uint64_t v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = on ? ((uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32)) : 0;
        while (k < KA) {
            const uint64_t q = C.m[k].q;
            const uint64_t zone = (q << __clzll(q)) - 1;
            uint32_t flags = 0, cnt = 0;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint64_t idx = c0 + 8 * tid + j;
                const bool ok = on && idx >= pos && v[j] * q <= zone;
                flags |= (uint32_t)ok << j;
                cnt += ok;
            }
            uint32_t excl;
            const uint32_t total = block_scan256(cnt, excl, sh);
            const uint32_t need = N - filled;
            uint32_t rank = excl;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (flags >> j & 1) {
                    if (rank < need) poly[2 * k][lds_ix(filled + rank)] = (uint32_t)__umul64hi(v[j], q);
                    if (rank == need - 1) s_last = 8 * tid + j;
                    rank++;
                }
            }
            __syncthreads();
            if (total >= need) {
                pos = c0 + s_last + 1;
                filled = 0;
                k++;
                __syncthreads();
                continue;
            }
            filled += total;
            pos = c0 + 8 * (uint64_t)active;
            break;
        }
        c0 += 8 * (uint64_t)active;
    }
    if (k < KA) {  // draw budget exhausted (never in practice): report, write nothing
        if (tid == 0) atomicOr(err, 1u);
        return;
    }
    __syncthreads();  // the a residues and the twiddles
    const uint32_t* b = bsrc + (size_t)blockIdx.x * KN;
    uint32_t* out = rows + (size_t)blockIdx.x * 2 * KN;
    // this thread's eight consecutive slots of poly ab (0: a, 1: b) of residue kk, Montgomery form
    auto store = [&](const uint32_t (&x)[8], int ab, uint32_t kk, const Mod& M) {
        uint32_t y[8];
        if constexpr (SGN) {
            // the scale c in Montgomery form: one signed product gives x c in (-q, q)
            const uint32_t cm = csub(shoup_mul(sc.c[kk], C.mont[kk], C.montp[kk], M.q), M.q);
#pragma unroll
            for (int e = 0; e < 8; e++) y[e] = sfix(smmul(x[e], cm, M), M.q);
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) y[e] = csub(shoup_mul(x[e], sc.c[kk], sc.cp[kk], M.q), M.q);
        }
        uint4* o = reinterpret_cast<uint4*>(out + (size_t)ab * KN + (size_t)kk * N + 8 * tid);
        o[0] = make_uint4(y[0], y[1], y[2], y[3]);
        o[1] = make_uint4(y[4], y[5], y[6], y[7]);
    };
    for (uint32_t kk = 0; kk < C.K; kk++) {
        const Mod M{C.m[kk].q, C.m[kk].q2, C.m[kk].qinv};
        if (kk < KA) {
            uint32_t x[2][8];  // a, b of residue kk at the coefficient positions tid + e N/8
#pragma unroll
            for (int e = 0; e < 8; e++) {
                x[0][e] = poly[2 * kk][lds_ix(tid + e * S::T)];
                x[1][e] = __builtin_nontemporal_load(b + (size_t)kk * N + tid + e * S::T);
            }
            // every thread holds its a values and is past the previous transform's
            // exchange reads before buffers kk, kk + 1 are written
            __syncthreads();
            if constexpr (SGN) ntt_digits<11, 2, true, 1>(x, tid, poly[kk], C.m[kk].twm, M, C.mont[kk]);
            else ntt_digits<11, 2, false>(x, tid, poly[kk], C.m[kk].twm, M);
            store(x[0], 0, kk, M);
            store(x[1], 1, kk, M);
        } else {  // b alone (a of this residue is not read)
            uint32_t x[1][8];
#pragma unroll
            for (int e = 0; e < 8; e++) x[0][e] = __builtin_nontemporal_load(b + (size_t)kk * N + tid + e * S::T);
            __syncthreads();  // past the previous transform's exchange reads of buffer kk
            if constexpr (SGN) ntt_digits<11, 1, true, 1>(x, tid, poly[kk], C.m[kk].twm, M, C.mont[kk]);
            else ntt_digits<11, 1, false>(x, tid, poly[kk], C.m[kk].twm, M);
            store(x[0], 1, kk, M);
        }
    }
}

// ---------------------------------------------------------------------------
// encode_database (encode_db.rs:106-164) on the device: entry e of the byte
// stream lands in shard e / N, local index e % N; column c is the LE u16 of
// bytes (2c, 2c+1), a lone last byte, or 0 (extract_column_value,
// encode_db.rs:169-181).  out: [S][cols][K][N] residues, zero past the last
// entry.  `bytes` holds the entries of shards first_shard.. of this slice.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_encode_entries(const uint8_t* __restrict__ bytes, uint64_t num_entries_in_slice, uint32_t entry_size,
                 uint32_t cols, uint32_t shards, uint32_t* __restrict__ out, CtxDev C) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)shards * cols * C.n) return;
    const uint32_t j = (uint32_t)(gid % C.n);
    const size_t sc = gid / C.n;                    // shard * cols + c
    const uint32_t c = (uint32_t)(sc % cols);
    const uint64_t e = (sc / cols) * (uint64_t)C.n + j;
    uint32_t v = 0;
    if (e < num_entries_in_slice) {
        const uint8_t* p = bytes + e * entry_size;
        const uint32_t bo = 2 * c;
        if (bo + 1 < entry_size) v = (uint32_t)p[bo] | ((uint32_t)p[bo + 1] << 8);
        else if (bo < entry_size) v = p[bo];
    }
    uint32_t* o = out + sc * C.K * C.n;
    for (uint32_t k = 0; k < C.K; k++) o[(size_t)k * C.n + j] = reduce32(v, C.m[k].q, C.onep[k]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int LOGN>
static hipError_t launch_ntt_t(const uint32_t* src, size_t ss, uint32_t* dst, size_t ds, size_t count,
                               const CtxDev& C, const ScaleDev& sc, bool inverse, hipStream_t st) {
    const size_t blocks = count * C.K;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    if (inverse)
        hipLaunchKernelGGL(k_ntt_inv<LOGN>, dim3((unsigned)blocks), dim3(NttShape<LOGN>::T), 0, st, src, ss, dst, ds, C,
                           sc);
    else
        hipLaunchKernelGGL(k_ntt_fwd<LOGN>, dim3((unsigned)blocks), dim3(NttShape<LOGN>::T), 0, st, src, ss, dst, ds, C,
                           sc);
    return hipGetLastError();
}

hipError_t launch_ntt_strided(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, size_t count,
                              const CtxDev& C, const ScaleDev& sc, bool inverse, hipStream_t st) {
    if (!inverse && C.ntt_mfma && ntt_mfma_supported(C))
        return launch_ntt_fwd_mfma(src, src_stride, dst, dst_stride, count, C.mfma, C, sc, st);
    switch (C.logn) {
        case 8: return launch_ntt_t<8>(src, src_stride, dst, dst_stride, count, C, sc, inverse, st);
        case 9: return launch_ntt_t<9>(src, src_stride, dst, dst_stride, count, C, sc, inverse, st);
        case 10: return launch_ntt_t<10>(src, src_stride, dst, dst_stride, count, C, sc, inverse, st);
        case 11: return launch_ntt_t<11>(src, src_stride, dst, dst_stride, count, C, sc, inverse, st);
        case 12: return launch_ntt_t<12>(src, src_stride, dst, dst_stride, count, C, sc, inverse, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt(uint32_t* data, size_t count, const CtxDev& C, const ScaleDev& sc, bool inverse,
                      hipStream_t st) {
    const size_t KN = (size_t)C.K * C.n;
    return launch_ntt_strided(data, KN, data, KN, count, C, sc, inverse, st);
}

static inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_db_digits(const uint32_t* in, uint32_t* out, size_t P, const CtxDev& C, hipStream_t st) {
    if (!P) return hipSuccess;
    hipLaunchKernelGGL(k_db_digits, dim3(blocks_for(P * C.n, 256)), dim3(256), 0, st, in, out, P, C);
    return hipGetLastError();
}

hipError_t launch_rlwe_digits(const uint32_t* in, uint32_t* out, size_t P, const CtxDev& C, hipStream_t st) {
    if (!P) return hipSuccess;
    hipLaunchKernelGGL(k_rlwe_digits, dim3(blocks_for(P * 2 * C.n, 256)), dim3(256), 0, st, in, out, P, C);
    return hipGetLastError();
}

hipError_t launch_first_dim(const uint64_t* qdig, const uint32_t* rows, uint32_t* out, uint32_t Q, uint32_t cols,
                            uint32_t out_stride, uint32_t out_off, const CtxDev& C, hipStream_t st) {
    const size_t total = (size_t)Q * cols * C.K * (C.n / 4);
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_first_dim, dim3(blocks_for(total, 256)), dim3(256), 0, st, qdig, rows, out, Q, cols,
                       out_stride, out_off, C);
    return hipGetLastError();
}

hipError_t launch_ext_mac(const uint32_t* dig, const uint32_t* rgsw, uint32_t* out, size_t P, const CtxDev& C,
                          hipStream_t st) {
    const size_t total = P * C.K * (C.n / 4);
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_ext_mac, dim3(blocks_for(total, 256)), dim3(256), 0, st, dig, rgsw, out, P, C);
    return hipGetLastError();
}

hipError_t launch_sum_cols(uint32_t* resp, uint32_t Q, uint32_t cols, const CtxDev& C, hipStream_t st) {
    const size_t total = (size_t)Q * 2 * C.K * C.n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_sum_cols, dim3(blocks_for(total, 256)), dim3(256), 0, st, resp, Q, cols, C);
    return hipGetLastError();
}

hipError_t launch_rows_seeded(const uint32_t* seeds, const uint32_t* bsrc, uint32_t* rows, size_t count,
                              uint32_t* err, const CtxDev& C, const ScaleDev& sc, hipStream_t st, LaunchEv ev,
                              uint32_t a_res, const uint32_t* a_full) {
    if (!count) return hipSuccess;
    if (C.logn != 11 || count > 0x7fffffffu) return hipErrorInvalidValue;
    if (a_res < 1 || a_res > C.K) a_res = C.K;
    // canonical inputs (below q) through 11 signed stages with one reduction after six
    bool sgn = true;
    for (uint32_t k = 0; k < C.K; k++) sgn = sgn && signed_full_ok(C.m[k].q, C.m[k].q - 1, 11, 6);
    if (sgn)
        hipExtLaunchKernelGGL(k_rows_seeded<true>, dim3((unsigned)count), dim3(256), 0, st, ev.start, ev.stop, 0,
                              seeds, bsrc, rows, err, C, sc, a_full, a_res, C.ell);
    else
        hipExtLaunchKernelGGL(k_rows_seeded<false>, dim3((unsigned)count), dim3(256), 0, st, ev.start, ev.stop, 0,
                              seeds, bsrc, rows, err, C, sc, a_full, a_res, C.ell);
    return hipGetLastError();
}

hipError_t launch_expand_seeds(const uint32_t* seeds, uint32_t* out, size_t count, size_t out_stride, uint32_t* err,
                               const CtxDev& C, hipStream_t st) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_expand_seeds, dim3((unsigned)count), dim3(256), 0, st, seeds, out, out_stride, err, C);
    return hipGetLastError();
}

hipError_t launch_encode_entries(const uint8_t* bytes, uint64_t num_entries, uint32_t entry_size, uint32_t cols,
                                 uint32_t shards, uint32_t* out, const CtxDev& C, hipStream_t st) {
    const size_t total = (size_t)shards * cols * C.n;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_encode_entries, dim3(blocks_for(total, 256)), dim3(256), 0, st, bytes, num_entries, entry_size,
                       cols, shards, out, C);
    return hipGetLastError();
}

}  // namespace inspire
