// InspiRING packing on the device (src/inspiring/inspiring2.rs, the partial
// packing used by respond_inspiring, src/pir/respond.rs:284-386).
#pragma once
#include <cstddef>
#include <cstdint>

#include "context.hpp"

namespace inspire {

// How a packed variant is answered.
enum BatchKind : uint8_t {
    KIND_PLAIN = 0,      // InsPIRe^0 (respond / respond_seeded)
    KIND_INSPIRING = 1,  // InspiRING packing (respond_inspiring, respond.rs:284-386)
    KIND_TREE = 2,       // tree packing (respond_one_packing, respond.rs:213-268)
};

// One runnable query of a respond batch (shared with respond.cpp).
struct BatchItem {
    uint32_t idx;   // position in the caller's batch
    uint64_t dig;   // device address of the shard's digits
    uint8_t kind;   // BatchKind
};

// Kernel-side description of a packing batch.
// Per-query roles of a launch that serves InsPIRe^0 beside the packed
// variants: role[q] 0 packs (InsPIRe^1/^2, block [2][K][N]), 1 answers with
// InsPIRe^0's combined and column ciphertexts (block [1 + gamma][2][K][N]);
// qout[q] is the word offset of query q's block in the output.  Null role:
// every query packs, query q's block at q 2KN.  Roles need the fused front
// end and the two-modulus chain (pack_mixed_supported).
struct PackRoles {
    const uint32_t* role = nullptr;
    const uint64_t* qout = nullptr;
};

struct PackArgs {
    const uint32_t* r_pre;       // [Q][gamma][K][N] step-1 sums before the automorphism
    const uint32_t* perm;        // [gamma][N] NTT-domain automorphism tables of g^i
    const uint32_t* w_all;       // [gamma-1][ell][K][N] Montgomery form
    const uint32_t* y;           // per query: NTT of y_body [ell][K][N] or of y_all [gamma-1][ell][K][N], Montgomery
    const uint64_t* y_off;       // [Q] word offset of each query's y block
    const uint32_t* y_mode;      // [Q] 0: rotate y_body by perm, 1: explicit y_all
    const uint32_t* y_count;     // [Q] rotations available (y_all.len(), or gamma-1 for y_body)
    uint32_t y_coeff;            // 1: y holds y_body in the coefficient domain (plain); the chain
                                 //    kernel transforms it (two-modulus kernel only)
    const uint64_t* bpart;       // [Q][nparts][gamma][K] partial sums of bhat (b_poly = n^-1 * sum)
    uint32_t nparts;
    uint32_t ninv0, ninv1;       // n^-1 mod q_k
    uint32_t ninvm[2], ninvmp[2];  // n^-1 * 2^32 mod q_k (Shoup pairs)
    uint32_t* out;               // [Q][2][K][N] packed ciphertext (a_hat, b)
    uint32_t gamma;
    uint32_t residue_wise;       // 1: corrected automorphism on every residue
    uint32_t q0r64, q0r64p;      // 2^64 mod q0 (Shoup pair): plain -> reference Montgomery value
    uint32_t q1r64i, q1r64ip;    // 2^-64 mod q1 (Shoup pair): reference value -> plain
    uint32_t one_q1p;            // Shoup companion of 1 mod q1
    uint32_t zpowm[2][4];        // z^j 2^32 mod q_k (gadget base z), j < ell
    uint32_t zinvm[2];           // z^-(ell-1) 2^32 mod q_k
    uint32_t zrecm[2][4];        // -z^(j-(ell-1)) 2^32 mod q_k, j < ell - 1
    // signed chain: R[i] enters its multiply-accumulate as r * rmul[k] (2^32 mod q_k,
    // or 2^-32 mod q1 for the rebuilt residue 1, whose value carries 2^64), S as S * 2^32
    uint32_t rmul[2], smul[2];
    PackRoles roles;             // per-query roles and output offsets (null role: all packed)
};

// Keys given as y_body alone are rotated in the NTT domain by the slot
// permutation of tau_{g^i}; generate_rotations (inspiring2.rs:1309-1323)
// applies the coefficient map of g^i, which is that permutation only for odd
// g^i.  Paths that rotate y_body on the device refuse even generators with
// this message (the host-buffer API computes those rotations as written).
constexpr const char* kEvenGeneratorMsg =
    "y_body keys rotated on the device need an odd InspiRING generator (gamma | 2d); "
    "send y_all or use the host-buffer API for this column count";

// Builds (once per gamma) the resident PackParams/OfflinePackingKeys tables.
PackTables& pack_tables(inspire_ctx* ctx, uint32_t gamma);
// The tables of gamma when already built, else nullptr (builds nothing).
PackTables* find_pack_tables(inspire_ctx* ctx, uint32_t gamma);

// Device scratch of a packing batch.  Null members are taken from the
// context's workspaces (grown on demand).
struct PackWs {
    uint32_t* rpre = nullptr;   // [Q][gamma][K][N] step-1 sums before the automorphism
    uint64_t* bpart = nullptr;  // [Q][nparts][gamma][K] b sums
    uint32_t* cols = nullptr;   // [Q][2 gamma][K][N] column ciphertexts (general front end only)
};
// Words of each PackWs part for Q queries (u32 words; bpart in u64 words).
struct PackWsSizes {
    size_t rpre = 0, bpart = 0, cols = 0;
};
PackWsSizes pack_ws_sizes(const inspire_ctx* ctx, const PackTables& T, uint32_t Q);

// Device-only packing of a staged group (rows in NTT/Montgomery form):
// out [Q][2][K][N].  y: NTT-domain Montgomery client keys with per-query
// offsets, modes (0: y_body rotated on the fly, 1: explicit y_all) and counts.
// d_qdig: per-query device addresses of the shard digits (databases of gamma columns).
// y_coeff: d_y holds coefficient-domain y_body (only when pack_chain_crt_supported).
void pack_run(inspire_ctx* ctx, uint32_t gamma, uint32_t Q, const uint32_t* rows, const uint64_t* d_qdig,
              const uint32_t* d_y, const uint64_t* d_yoff, const uint32_t* d_ymode, const uint32_t* d_ycount,
              uint32_t* d_out, bool y_coeff, const PackWs& ws = PackWs{}, const PackRoles& roles = PackRoles{});
// True when one pack_run launch sequence can also answer InsPIRe^0 queries
// (PackRoles) for databases of gamma columns: the fused front end applies
// (gamma = 2^k, gamma^2 | 2d) and the two-modulus chain kernel runs.
bool pack_mixed_supported(inspire_ctx* ctx, uint32_t gamma);
// Residues of the RGSW rows' a that packing gamma columns reads: 1 when the
// fused front end runs in the reference's residue handling with two moduli
// (only residue 0 of step 1 is kept, and only its a enters), K otherwise.
uint32_t pack_rows_a_residues(inspire_ctx* ctx, uint32_t gamma);
// True when the packing chain transforms coefficient-domain y_body itself;
// otherwise pack_run_device_ybody needs y_ws for the transformed keys.
bool pack_ybody_in_chain(const inspire_ctx* ctx);
// Device-resident y_body [Q][ell][K][N] (coefficient domain).  y_ws: [Q][ell][K][N]
// scratch for the transformed keys when the chain does not transform them
// (nullptr: a context workspace).
void pack_run_device_ybody(inspire_ctx* ctx, uint32_t gamma, uint32_t Q, const uint32_t* rows,
                           const uint64_t* d_qdig, const uint32_t* y_body, uint32_t* d_out,
                           const PackWs& ws = PackWs{}, uint32_t* y_ws = nullptr);
// Host-side keys of the queries `items` (y_body or y_all), responses copied back.
void pack_respond_host(inspire_ctx* ctx, uint32_t gamma, const BatchItem* items, uint32_t Q, const uint32_t* rows,
                       const uint64_t* d_qdig, const uint32_t* d_err, const inspire_query* qs, inspire_response* rs,
                       int32_t* status);

hipError_t launch_pack_b0(const uint32_t* cols, uint64_t* bpart, uint32_t Q, uint32_t gamma, const CtxDev& C,
                          hipStream_t st);
// Fused first dimension + step 1 + b sums for gamma in {1, 2, ..., 64} with gamma^2 | 2n.
// qdig[q]: device address of query q's shard digits (a valid shard also for a missing one).
hipError_t launch_pack_front(const uint64_t* qdig, const uint32_t* rows, uint32_t* r_pre, uint64_t* bpart,
                             const uint32_t* twist, const uint2* dftw, uint32_t Q, uint32_t gamma, bool rw,
                             const CtxDev& C, hipStream_t st, LaunchEv ev = LaunchEv{},
                             const PackRoles& roles = PackRoles{}, uint32_t* out = nullptr);
// stepu: [gamma][K][N] u = w_s^(g^(n-i)) in Montgomery form; ginvm: gamma^-1 2^32 mod q_0, q_1.
hipError_t launch_pack_step1(const uint32_t* cols, const uint32_t* stepu, uint32_t* r_pre, uint32_t Q,
                             uint32_t gamma, uint32_t kres, uint2 ginvm, const CtxDev& C, hipStream_t st);
// Two-modulus packing chain with every thread on both residues
// (pack_chain.hip), for d <= 2048 and ell <= 3; launch_pack_chain_scaled
// uses it whenever pack_chain_crt_supported holds.
bool pack_chain_crt_supported(const CtxDev& C);
hipError_t launch_pack_chain_crt(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                 hipStream_t st, LaunchEv ev = LaunchEv{});
// inv: inverse-NTT scaling (plain inverse: n^-1 folded into the last stage).
hipError_t launch_pack_chain_scaled(const PackArgs& A, uint32_t Q, const CtxDev& C, const ScaleDev& inv,
                                    hipStream_t st, LaunchEv ev = LaunchEv{});

}  // namespace inspire
