// Host driver of tree packing (respond_one_packing / respond_seeded_packed,
// src/pir/respond.rs:213-268, 428-435; pack_lwes,
// src/inspiring/automorph_pack.rs:418-487).
//
// Launch plan for m = min(cols, N) columns, L = log2 N levels:
//   k_tree_leaves   Q * m workgroups: first dimension of each column, its LWE,
//                   and levels 1 .. L - ceil(log2 m) (no node has an odd child)
//   k_tree_combine  one launch per remaining level, Q * min(m, 2^(L - l))
//                   workgroups; the root launch writes the packed ciphertext
// A single column is not tree-packed (automorph_pack.rs:436-443): the leaves
// launch writes it directly.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "modarith.hpp"
#include "respond.hpp"
#include "tree.hpp"

namespace inspire {

static uint32_t bitrev32(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

void set_galois_keys(inspire_ctx* ctx, const uint64_t* keys, uint32_t count) {
    if (count != ctx->logn)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "galois keys: expected log2(d) = " + std::to_string(ctx->logn) +
                                                      " key-switching matrices (setup.rs:167-173), got " +
                                                      std::to_string(count)};
    if (!tree_supported(ctx->dev))
        throw Error{INSPIRE_ERR_UNSUPPORTED, "tree packing needs d in [256, 2048], gadget_len <= 3 on the GPU path"};
    const size_t KN = (size_t)ctx->K * ctx->N, words = (size_t)count * ctx->ell * 2 * KN;
    for (size_t w = 0; w < words; w++)
        if (keys[w] >= ctx->hm[(w / ctx->N) % ctx->K].q)
            throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "galois keys: a residue is not reduced"};
    DevBuf gk(words * 4);
    upload_u64(ctx, keys, words, gk.get<uint32_t>());
    uint64_t mont[kMaxK];
    for (uint32_t k = 0; k < ctx->K; k++) mont[k] = (1ull << 32) % ctx->hm[k].q;
    hip_check(launch_ntt(gk.get<uint32_t>(), (size_t)count * ctx->ell * 2, ctx->dev, scale_fwd(ctx, mont), false,
                         ctx->stream),
              "galois keys ntt");
    // y = X^(N >> l) of level l in the NTT domain: slot p holds
    // psi^((N >> l) (2 br(p) + 1)), Montgomery form
    std::vector<uint32_t> yh((size_t)ctx->logn * KN);
    for (uint32_t l = 1; l <= ctx->logn; l++)
        for (uint32_t k = 0; k < ctx->K; k++) {
            const uint64_t q = ctx->hm[k].q;
            const uint64_t r32 = (1ull << 32) % q;
            for (uint32_t p = 0; p < ctx->N; p++) {
                const uint64_t e = ((uint64_t)(ctx->N >> l) * (2ull * bitrev32(p, ctx->logn) + 1)) % (2ull * ctx->N);
                yh[((size_t)(l - 1) * ctx->K + k) * ctx->N + p] = (uint32_t)mulmod64(powmod64(ctx->hm[k].psi, e, q), r32, q);
            }
        }
    DevBuf yd(yh.size() * 4);
    hip_check(hipMemcpyAsync(yd.get(), yh.data(), yh.size() * 4, hipMemcpyHostToDevice, ctx->stream), "yhat");
    hip_check(hipStreamSynchronize(ctx->stream), "galois keys");
    ctx->galois = std::move(gk);
    ctx->yhat = std::move(yd);
    ctx->galois_count = count;
}

uint32_t tree_b_dim(const inspire_ctx* ctx, uint32_t cols) {
    const bool rw = (root(ctx)->params.flags & INSPIRE_FLAG_RESIDUE_WISE_AUTOMORPHISM) != 0;
    return (ctx->K == 2 && !rw && cols >= 2) ? 2 * ctx->N : ctx->N;
}

size_t tree_out_words(const inspire_ctx* ctx, uint32_t cols) {
    return (size_t)ctx->K * ctx->N + (size_t)ctx->K * tree_b_dim(ctx, cols);
}

void tree_run(inspire_ctx* ctx, uint32_t cols, uint32_t Q, const uint32_t* rows, const uint64_t* d_qdig,
              uint32_t* out) {
    const inspire_ctx* R = root(ctx);
    if (!R->galois_count)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT,
                    "tree packing needs the CRS galois keys (inspire_ctx_set_galois_keys)"};
    if (!Q || !cols) return;
    const uint32_t N = ctx->N, K = ctx->K, L = ctx->logn;
    const uint32_t m = cols < N ? cols : N;
    const size_t KN = (size_t)K * N;
    TreeArgs A{};
    A.qdig = d_qdig;
    A.rows = rows;
    A.ks = R->galois.get<uint32_t>();
    A.yhat = R->yhat.get<uint32_t>();
    A.m = m;
    A.logn = L;
    A.ref_layout = tree_b_dim(ctx, cols) == 2 * N;
    A.out = out;
    A.out_words = (uint32_t)tree_out_words(ctx, cols);
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t q = ctx->hm[k].q;
        const uint64_t ninv = powmod64(N, q - 2, q);
        A.ninvm[k] = (uint32_t)mulmod64(ninv, (1ull << 32) % q, q);
        A.ninvmp[k] = shoup_pre(A.ninvm[k], (uint32_t)q);
        A.dmod[k] = (uint32_t)(N % q);
        A.dmodp[k] = shoup_pre(A.dmod[k], (uint32_t)q);
        // gadget powers for the recovered last digit, Montgomery form
        const uint64_t z = (1ull << ctx->dev.gadget_log) % q, r32 = (1ull << 32) % q;
        uint64_t zj = 1;
        for (uint32_t j = 0; j < ctx->ell && j < 4; j++) {
            A.zpowm[k][j] = (uint32_t)mulmod64(zj, r32, q);
            if (j + 1 < ctx->ell) zj = mulmod64(zj, z, q);
        }
        const uint64_t zinv = powmod64(zj, q - 2, q);  // zj = z^(ell-1)
        A.zinvm[k] = (uint32_t)mulmod64(zinv, r32, q);
        // the recovered digit as one sum: z^-(ell-1) R + sum_j -z^(j-(ell-1)) NTT(D_j)
        uint64_t zc = zinv;
        for (uint32_t j = 0; j + 1 < ctx->ell && j < 4; j++) {
            const uint64_t c = mulmod64(zc, r32, q);
            A.zrecm[k][j] = (uint32_t)(c ? q - c : 0);
            zc = mulmod64(zc, z, q);
        }
    }
    const size_t node_words = (size_t)Q * m * 3 * KN, lweb_words = (size_t)Q * m * K;
    ctx->ws3.reserve((node_words + lweb_words) * 4);
    A.node = ctx->ws3.get<uint32_t>();
    A.lweb = A.node + node_words;
    const ScaleDev inv = scale_inv(ctx, nullptr);
    // levels without an odd child anywhere: 2^(L - l) >= m
    uint32_t lm = 0;
    while ((1u << lm) < m) lm++;
    const uint32_t top = L - lm;  // last such level (0 when m = N)
    {
        ProfScope ps(ctx, P_TREE_LEAVES);
        A.level_lo = 1;
        A.level_hi = m == 1 ? 0 : top;
        hip_check(launch_tree(A, true, Q * m, ctx->dev, inv, ctx->stream), "tree leaves");
    }
    if (m == 1) return;
    ProfScope ps(ctx, P_TREE_COMBINE);
    for (uint32_t l = top + 1; l <= L; l++) {
        A.level_lo = A.level_hi = l;
        const uint32_t half = 1u << (L - l);
        hip_check(launch_tree(A, false, Q * (m < half ? m : half), ctx->dev, inv, ctx->stream), "tree combine");
    }
}

}  // namespace inspire
