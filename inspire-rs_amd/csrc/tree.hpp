// Tree packing (respond_one_packing, src/pir/respond.rs:213-268;
// src/inspiring/automorph_pack.rs) on the device.
#pragma once
#include <cstddef>
#include <cstdint>

#include "kernels.hpp"

struct inspire_ctx;

namespace inspire {

// Kernel-side description of a tree-packing batch (tree_kernels.hip).
struct TreeArgs {
    const uint64_t* qdig;   // [Q] device address of each query's shard digits ([cols][ell][K][N])
    const uint32_t* rows;   // [Q][ell][2][K][N] RGSW rows ell..2ell-1, NTT domain, Montgomery form
    const uint32_t* ks;     // [log N][ell][2][K][N] galois keys (ServerCrs.galois_keys), NTT, Montgomery
    const uint32_t* yhat;   // [log N][K][N] NTT(X^(N >> l)) of level l at l - 1, Montgomery form
    uint32_t* node;         // [Q][m][3][K][N] tree nodes: a (coefficient domain), a^ and b (NTT domain)
    uint32_t* lweb;         // [Q][m][K] the LWE b of each column (coefficient 0 of its b)
    uint32_t* out;          // [Q][out_words] a [K][N], then b [K][b_dim]
    uint32_t m;             // LWEs packed (columns, at most N)
    uint32_t logn;
    uint32_t level_lo, level_hi;  // leaves: levels run after the first dimension; combine: the level
    uint32_t ref_layout;    // 1: the reference's two-modulus b (b_dim = 2N), else b_dim = N
    uint32_t out_words;
    uint32_t ninvm[2], ninvmp[2];  // n^-1 2^32 mod q_k (Shoup pairs)
    uint32_t dmod[2], dmodp[2];    // N mod q_k (Shoup pairs)
    uint32_t zpowm[2][4];          // z^j 2^32 mod q_k (gadget base z), j < ell
    uint32_t zinvm[2];             // z^-(ell-1) 2^32 mod q_k
    uint32_t zrecm[2][4];          // -z^(j-(ell-1)) 2^32 mod q_k, j < ell - 1
};

bool tree_supported(const CtxDev& C);
// leaves: k_tree_leaves over Q * m workgroups; else k_tree_combine for level
// A.level_lo over `blocks` = Q * min(m, 2^(log N - level)) workgroups.
hipError_t launch_tree(const TreeArgs& A, bool leaves, uint32_t blocks, const CtxDev& C, const ScaleDev& inv,
                       hipStream_t st);

// b dimension of a tree-packed response with `cols` columns: the reference
// rebuilds b from its 2N residue words when there are two moduli and at least
// two LWEs (automorph_pack.rs:473-484); N otherwise and in the corrected mode.
uint32_t tree_b_dim(const inspire_ctx* ctx, uint32_t cols);
// Words of one tree-packed response: a [K][N] and b [K][b_dim].
size_t tree_out_words(const inspire_ctx* ctx, uint32_t cols);

// Tree-packs a staged group: rows [Q][ell][2][K][N] (NTT, Montgomery form),
// d_qdig per-query shard digits of databases with `cols` columns -> out
// [Q][tree_out_words].  Needs the context's galois keys.
void tree_run(inspire_ctx* ctx, uint32_t cols, uint32_t Q, const uint32_t* rows, const uint64_t* d_qdig,
              uint32_t* out);

// Uploads ServerCrs.galois_keys (setup.rs:167-173): count = log2 N
// KeySwitchingMatrix's of ell rows (a, b), [count][ell][2][K][N] u64.
void set_galois_keys(inspire_ctx* ctx, const uint64_t* keys, uint32_t count);

}  // namespace inspire
