// Host-visible kernel interface: constant bundles passed by value and launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ntt.hpp"

namespace inspire {

constexpr int kMaxK = 2;      // CRT moduli
constexpr int kMaxEll = 4;    // gadget digits

// Raw buffer view of a wave-uniform base address (gfx9 descriptor: stride 0,
// 32-bit records, 2^31 - 1 bytes).  A load or store takes the lane's byte
// offset in one VGPR and the uniform word offset in an SGPR, so a kernel that
// reads many arrays at uniform strides spends no vector instructions on
// addresses; through a pointer read from memory (a query's shard address)
// the compiler would instead form a 64-bit address per lane and per load.
struct BufU32 {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ explicit BufU32(const void* base)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7fffffff, 0x00020000)) {}
    // word `off` (uniform) past the base, plus the lane's byte offset vb
    __device__ __forceinline__ uint32_t ld(uint32_t vb, uint32_t off) const {
        return __builtin_amdgcn_raw_buffer_load_b32(r, vb, off * 4u, 0);
    }
    // the same, non-temporal (data read once: kept out of the caches' way)
    __device__ __forceinline__ uint32_t ldnt(uint32_t vb, uint32_t off) const {
        return __builtin_amdgcn_raw_buffer_load_b32(r, vb, off * 4u, 2);
    }
    __device__ __forceinline__ void st(uint32_t vb, uint32_t off, uint32_t v) const {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, vb, off * 4u, 0);
    }
};

// Constant tables of the matrix-core forward NTT at d = 2048 (ntt_mfma.hip).
struct MfmaNttTables {
    const uint32_t* tw1;  // [K][2048] psi^j 2^32 mod q (the negacyclic twist, Montgomery form)
    const uint32_t* tw2;  // [K][32][64] w^(j1 m2) 2^32 mod q, w = psi^2
    const int8_t* w64;    // [K][4][64][64] 7-bit limbs of w^(32 j2 m2)
    const int8_t* w32;    // [K][4][32][32] 7-bit limbs of w^(64 m1 j1)
    const uint32_t* c1;   // [K][8] 2^(7s) 2^32 mod q, s < 7 (limb-sum recombination)
};

// Everything a kernel needs about the parameter set.
struct CtxDev {
    ModDev m[kMaxK];
    uint32_t K, logn, n, ell;
    uint32_t gadget_log;          // log2(gadget_base), at most 32
    uint32_t gadget_mask;         // gadget_base - 1
    uint32_t q0_inv_q1, q0_inv_q1p;  // CRT: q0^-1 mod q1 (Shoup pair, mod q1)
    uint32_t one_q1p;                // Shoup companion of 1 mod q1 (x mod q1 for x < 2^32)
    uint32_t onep[kMaxK];            // Shoup companion of 1 mod q_k (reduce32)
    uint32_t small_digits;           // gadget_base <= every q_k: digits need no reduction
    uint64_t q;                   // composite modulus
    uint32_t delta[kMaxK], deltap[kMaxK];  // Delta mod q_k (Shoup pairs)
    uint32_t mont[kMaxK], montp[kMaxK];    // 2^32 mod q_k (Montgomery-form conversion)
    MfmaNttTables mfma;           // null when the matrix-core transform does not apply
    uint32_t ntt_mfma;            // 1: forward d = 2048 transforms run on the matrix cores
    uint32_t chain_v;             // values per thread of the two-modulus packing chain: 4 (0) or 8
};

// d = 2048, every modulus below 2^28 and the tables built.
bool ntt_mfma_supported(const CtxDev& C);

// Per-residue output scaling.  Forward: x -> c x when enable.  Inverse: the
// last stage computes c (X + Y) and cw (X - Y).
struct ScaleDev {
    uint32_t enable;
    uint32_t c[kMaxK], cp[kMaxK], cw[kMaxK], cwp[kMaxK];
};

// Events a launch records around its kernel's execution (hipExtLaunchKernelGGL):
// the start and stop of the dispatch itself, with no extra stream operations.
// Null events: a plain launch.
struct LaunchEv {
    hipEvent_t start = nullptr, stop = nullptr;
};

// In-place transform of count contiguous [K][N] polys.
hipError_t launch_ntt(uint32_t* data, size_t count, const CtxDev& C, const ScaleDev& sc, bool inverse,
                      hipStream_t st);
// Transform poly p from src + p*src_stride into dst + p*dst_stride (strides in words).
hipError_t launch_ntt_strided(const uint32_t* src, size_t src_stride, uint32_t* dst, size_t dst_stride, size_t count,
                              const CtxDev& C, const ScaleDev& sc, bool inverse, hipStream_t st);
hipError_t launch_encode_entries(const uint8_t* bytes, uint64_t num_entries, uint32_t entry_size, uint32_t cols,
                                 uint32_t shards, uint32_t* out, const CtxDev& C, hipStream_t st);
hipError_t launch_db_digits(const uint32_t* in, uint32_t* out, size_t P, const CtxDev& C, hipStream_t st);
hipError_t launch_rlwe_digits(const uint32_t* in, uint32_t* out, size_t P, const CtxDev& C, hipStream_t st);
// qdig[q]: device address of query q's shard digits ([cols][ell][K][N]), 0 for a missing shard.
hipError_t launch_first_dim(const uint64_t* qdig, const uint32_t* rows, uint32_t* out, uint32_t Q, uint32_t cols,
                            uint32_t out_stride, uint32_t out_off, const CtxDev& C, hipStream_t st);
hipError_t launch_ext_mac(const uint32_t* dig, const uint32_t* rgsw, uint32_t* out, size_t P, const CtxDev& C,
                          hipStream_t st);
// seeds [count][8] u32 key words; seed i expands into out + i*out_stride ([K][N]).
hipError_t launch_expand_seeds(const uint32_t* seeds, uint32_t* out, size_t count, size_t out_stride, uint32_t* err,
                               const CtxDev& C, hipStream_t st);
hipError_t launch_sum_cols(uint32_t* resp, uint32_t Q, uint32_t cols, const CtxDev& C, hipStream_t st);
// d = 2048: seeded rows expanded and transformed in one pass.  Row i: seed
// seeds + 8i, b at bsrc + i*K*N; out rows + i*2*K*N (a, then b), Montgomery
// form via sc.  Residues of a: a_res (1..K) for every row, or K for the rows
// of the queries q (row i of query i / ell) with a_full[q] != 0; the a of the
// other residues is not written.  Returns hipErrorInvalidValue for other degrees.
hipError_t launch_rows_seeded(const uint32_t* seeds, const uint32_t* bsrc, uint32_t* rows, size_t count,
                              uint32_t* err, const CtxDev& C, const ScaleDev& sc, hipStream_t st,
                              LaunchEv ev = LaunchEv{}, uint32_t a_res = kMaxK, const uint32_t* a_full = nullptr);

// Issue-rate probes (rates.hip): kind 0 the Montgomery product, 1 the signed
// Montgomery product, 2 the 64-bit multiply-add; out holds blocks * 256 words.
hipError_t launch_rate_probe(int kind, uint32_t* out, uint32_t blocks, uint32_t iters, const ModDev& m,
                             hipStream_t st);
uint64_t rate_ops_per_launch(uint32_t blocks, uint32_t iters);
// Streaming read of `bytes` (a multiple of 16) at src; out holds blocks * 256 words.
hipError_t launch_read_probe(const void* src, size_t bytes, uint32_t* out, uint32_t blocks, hipStream_t st);

}  // namespace inspire
