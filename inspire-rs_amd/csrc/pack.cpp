// InspiRING tables and the packed-respond driver (respond_inspiring,
// src/pir/respond.rs:284-386).
//
// Everything query-independent that the reference rebuilds per query is built
// once per (context, gamma) and kept resident:
//   PackParams::new             (inspiring2.rs:107-182): generator, g^i, the
//                               NTT-domain automorphism tables, gamma^-1 and
//                               the NTT-form monomials, folded here into the
//                               step-1 table gamma^-1 * NTT(X^(j g^(n-i)));
//   OfflinePackingKeys::generate (inspiring2.rs:448-506): w_mask from w_seed
//                               (ChaCha20, inspiring2.rs:816-835) and its
//                               rotations w_all_ntt.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "host_par.hpp"
#include "kernels.hpp"
#include "modarith.hpp"
#include "pack.hpp"
#include "respond.hpp"

namespace inspire {

namespace {

uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

// ChaCha20 block, 64-bit counter in words 12-13, stream 0 (rand_chacha 0.3).
void chacha_block(const uint32_t key[8], uint64_t ctr, uint32_t out[16]) {
    uint32_t x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                      key[4], key[5], key[6], key[7], (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u};
    uint32_t s[16];
    std::memcpy(s, x, sizeof(s));
    auto rotl = [](uint32_t v, int n) { return (v << n) | (v >> (32 - n)); };
    auto qr = [&](int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
    };
    for (int r = 0; r < 10; r++) {
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

// ChaCha20Rng::next_u64 stream: u64 t = word[2t] | word[2t+1] << 32.
struct ChaChaStream {
    uint32_t key[8];
    uint64_t block = 0;
    uint32_t buf[16];
    int idx = 16;
    explicit ChaChaStream(const uint8_t seed[32]) {
        for (int i = 0; i < 8; i++)
            key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) | ((uint32_t)seed[4 * i + 2] << 16) |
                     ((uint32_t)seed[4 * i + 3] << 24);
    }
    uint64_t next_u64() {
        if (idx >= 16) { chacha_block(key, block++, buf); idx = 0; }
        uint64_t v = (uint64_t)buf[idx] | ((uint64_t)buf[idx + 1] << 32);
        idx += 2;
        return v;
    }
};

// apply_automorphism (galois.rs:22-55) on composite coefficients, exactly as
// written: also covers even exponents, where the map is not a permutation.
std::vector<uint64_t> coeff_automorphism(const inspire_ctx* ctx, const uint64_t* poly, uint64_t g) {
    const uint32_t d = ctx->N;
    const uint64_t q = ctx->q;
    std::vector<uint64_t> res(d, 0), outp((size_t)ctx->K * d);
    const uint64_t q0 = ctx->hm[0].q;
    for (uint32_t i = 0; i < d; i++) {
        uint64_t c = poly[i];
        if (ctx->K == 2) {  // Poly::coeff (poly.rs:359-372)
            const uint64_t q1 = ctx->hm[1].q;
            const uint64_t a1 = poly[d + i], a0m = c % q1;
            const uint64_t diff = a1 >= a0m ? a1 - a0m : a1 + q1 - a0m;
            c = c + q0 * mulmod64(diff, powmod64(q0 % q1, q1 - 2, q1), q1);
            c %= q;
        }
        if (c == 0) continue;
        const uint64_t ni = (g * i) % (2ull * d);
        if (ni < d) res[ni] = (uint64_t)(((unsigned __int128)res[ni] + c) % q);
        else {
            const uint64_t ai = ni - d;
            res[ai] = res[ai] >= c ? res[ai] - c : q - c + res[ai];
        }
    }
    for (uint32_t k = 0; k < ctx->K; k++)
        for (uint32_t i = 0; i < d; i++) outp[(size_t)k * d + i] = res[i] % ctx->hm[k].q;
    return outp;
}

uint64_t mod_inverse_u64(uint64_t a, uint64_t m) {  // inspiring2.rs:1633-1640
    __int128 r0 = a, r1 = m, s0 = 1, s1 = 0;
    while (r1 != 0) {
        __int128 qq = r0 / r1, t;
        t = r1; r1 = r0 - qq * r1; r0 = t;
        t = s1; s1 = s0 - qq * s1; s0 = t;
    }
    if (r0 != 1) throw Error{INSPIRE_ERR_UNSUPPORTED, "num_to_pack must be invertible mod q"};
    __int128 x = s0 % (__int128)m;
    if (x < 0) x += m;
    return (uint64_t)x;
}

struct GenInfo {
    uint32_t generator;
    std::vector<uint32_t> gen_pows;  // [n]
};

GenInfo gen_info(uint32_t n, uint32_t gamma) {  // inspiring2.rs:114-128
    GenInfo g;
    g.generator = gamma < n ? (2 * n / gamma) + 1 : 5;
    g.gen_pows.resize(n);
    uint64_t v = 1;
    for (uint32_t i = 0; i < n; i++) { g.gen_pows[i] = (uint32_t)v; v = (v * g.generator) % (2ull * n); }
    return g;
}

}  // namespace

// Copies a table to the device on the calling context's stream (a lane may
// build tables while another lane captures a graph, and this runtime refuses
// null-stream copies while any capture is open).
static void upload_table(inspire_ctx* ctx, const DevBuf& dst, const void* src, size_t bytes, const char* what) {
    hip_check(hipMemcpyAsync(dst.get(), src, bytes, hipMemcpyHostToDevice, ctx->stream), what);
    hip_check(hipStreamSynchronize(ctx->stream), what);
}

PackTables& pack_tables(inspire_ctx* ctx, uint32_t gamma) {
    // The map is shared by the owner and its lanes: lookups and the first
    // build of a gamma's tables happen under its own lock.  Entries are never
    // erased, so references stay valid while other gammas are added.
    std::lock_guard<std::mutex> lk(root(ctx)->packs_mu);
    auto& packs = root(ctx)->packs;
    auto it = packs.find(gamma);
    if (it != packs.end()) return *it->second;
    const uint32_t n = ctx->N, K = ctx->K, ell = ctx->ell, logn = ctx->logn;
    const size_t KN = (size_t)K * n;
    // PackParams::new keeps n generator powers and packing_offline indexes them
    // by i < num_to_pack (inspiring2.rs:122-127, 926-1041): gamma <= n
    if (gamma == 0 || gamma > n)
        throw Error{INSPIRE_ERR_UNSUPPORTED, "InspiRING packs between 1 and d column LWEs (got " +
                                                 std::to_string(gamma) + " columns at d = " + std::to_string(n) + ")"};
    std::unique_ptr<PackTables> T(new PackTables());
    T->gamma = gamma;
    const GenInfo G = gen_info(n, gamma);
    // automorphism tables, get_automorph_table(t) = tables[(t-1)/2] (inspiring2.rs:186-198)
    std::vector<uint32_t> perm((size_t)gamma * n);
    for (uint32_t i = 0; i < gamma; i++) {
        const uint64_t t = 2ull * ((G.gen_pows[i] - 1) / 2) + 1;
        for (uint32_t s = 0; s < n; s++) {
            const uint64_t e = (t * (2ull * bitrev(s, logn) + 1)) % (2ull * n);
            perm[(size_t)i * n + s] = bitrev((uint32_t)((e - 1) / 2), logn);
        }
    }
    const uint64_t ginv = mod_inverse_u64(gamma, ctx->q);
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t q = ctx->hm[k].q;
        T->ginvm[k] = (uint32_t)mulmod64(ginv % q, (1ull << 32) % q, q);
    }
    // w_mask (inspiring2.rs:816-835) -> NTT on the device -> rotations with the
    // reference's residue handling -> Montgomery form.
    std::vector<uint64_t> mask((size_t)ell * KN);
    {
        ChaChaStream rng(ctx->w_seed);
        for (uint32_t d = 0; d < ell; d++)
            for (uint32_t i = 0; i < n; i++) {
                const uint64_t v = rng.next_u64() % ctx->q;
                for (uint32_t k = 0; k < K; k++) mask[(size_t)d * KN + (size_t)k * n + i] = v % ctx->hm[k].q;
            }
    }
    DevBuf tmp((size_t)ell * KN * 4);
    upload_u64(ctx, mask.data(), (size_t)ell * KN, tmp.get<uint32_t>());
    hip_check(launch_ntt(tmp.get<uint32_t>(), ell, ctx->dev, scale_none(), false, ctx->stream), "mask ntt");
    std::vector<uint64_t> mask_ntt((size_t)ell * KN);
    download_u64(ctx, tmp.get<uint32_t>(), (size_t)ell * KN, mask_ntt.data());
    const uint32_t nrot = gamma > 0 ? gamma - 1 : 0;
    std::vector<uint32_t> w_all((size_t)std::max(nrot, 1u) * ell * KN, 0);
    const bool rw = (ctx->params.flags & INSPIRE_FLAG_RESIDUE_WISE_AUTOMORPHISM) != 0;
    const uint64_t q0 = ctx->hm[0].q;
    const uint64_t r64_q0 = pow2_64_mod(q0);
    for (uint32_t i = 0; i < nrot; i++)
        for (uint32_t d = 0; d < ell; d++)
            for (uint32_t s = 0; s < n; s++) {
                const uint32_t src = perm[(size_t)i * n + s];
                const uint64_t v0 = mask_ntt[(size_t)d * KN + src];
                uint32_t* dst = w_all.data() + ((size_t)i * ell + d) * KN;
                dst[s] = (uint32_t)mulmod64(v0, (1ull << 32) % q0, q0);
                if (K == 2) {
                    const uint64_t q1 = ctx->hm[1].q;
                    uint64_t v1;
                    if (rw) v1 = mask_ntt[(size_t)d * KN + n + src];
                    else {  // apply_automorphism_ntt: residue 1 := (v0 * 2^64 mod q0) mod q1, a 2^64 Montgomery value
                        const uint64_t ref = mulmod64(v0, r64_q0, q0) % q1;
                        v1 = mulmod64(ref, powmod64(pow2_64_mod(q1), q1 - 2, q1), q1);
                    }
                    dst[n + s] = (uint32_t)mulmod64(v1, (1ull << 32) % q1, q1);
                }
            }
    // fused front end: gamma = 2^LG with gamma^2 | 2n (then g^i = 1 + (2n/gamma) i mod 2n)
    {
        uint32_t lg = 0;
        while ((1u << lg) < gamma) lg++;
        const bool pow2 = (1u << lg) == gamma && gamma <= 64;
        T->fast_front = pow2 && ((2ull * n) % ((uint64_t)gamma * gamma) == 0) && n >= 256;
        if (T->fast_front) {
            const uint32_t G = gamma, H = std::max(G / 2, 1u);
            std::vector<uint32_t> twist((size_t)K * G * n);
            std::vector<uint32_t> dftw((size_t)K * H * H * 2, 0);
            for (uint32_t k = 0; k < K; k++) {
                const uint64_t q = ctx->hm[k].q, psi = ctx->hm[k].psi;
                const uint64_t scale = mulmod64(ginv % q, (1ull << 32) % q, q);  // gamma^-1, Montgomery form
                for (uint32_t s = 0; s < n; s++) {
                    const uint64_t w = powmod64(psi, 2ull * bitrev(s, logn) + 1, q);
                    uint64_t wj = scale;
                    for (uint32_t j = 0; j < G; j++) {
                        twist[((size_t)k * G + j) * n + s] = (uint32_t)wj;
                        wj = mulmod64(wj, w, q);
                    }
                }
                // class c = 2*ci + 1: W_c = psi^(-c * 2n/G) (exponent mod 2n), powers W_c^m for m < G/2
                for (uint32_t ci = 0; ci < H; ci++) {
                    const uint64_t e = (2ull * n - ((2ull * ci + 1) * (2ull * n / G)) % (2ull * n)) % (2ull * n);
                    const uint64_t Wc = powmod64(psi, e, q);
                    uint64_t wm = 1;
                    for (uint32_t m = 0; m < H; m++) {
                        const size_t o = (((size_t)k * H + ci) * H + m) * 2;
                        dftw[o] = (uint32_t)wm;
                        dftw[o + 1] = shoup_pre((uint32_t)wm, (uint32_t)q);
                        wm = mulmod64(wm, Wc, q);
                    }
                }
            }
            T->twist.alloc(twist.size() * 4);
            T->dftw.alloc(dftw.size() * 4);
            upload_table(ctx, T->twist, twist.data(), twist.size() * 4, "twist");
            upload_table(ctx, T->dftw, dftw.data(), dftw.size() * 4, "dftw");
        }
    }
    if (!T->fast_front) {
        // step 1 by Horner's rule: u[i][k][s] = w_s^(g^((n-i) mod n) mod 2n), Montgomery form
        // (NTT(X^e)[s] = w_s^e, and e_ij = j g^(n-i) mod 2n is j times that exponent mod 2n)
        std::vector<uint32_t> stepu((size_t)gamma * KN);
        for (uint32_t k = 0; k < K; k++) {
            const uint64_t q = ctx->hm[k].q, psi = ctx->hm[k].psi, r32 = (1ull << 32) % q;
            for (uint32_t s = 0; s < n; s++) {
                const uint64_t w = powmod64(psi, 2ull * bitrev(s, logn) + 1, q);  // NTT slot s evaluates at w
                for (uint32_t i = 0; i < gamma; i++)
                    stepu[((size_t)i * K + k) * n + s] =
                        (uint32_t)mulmod64(powmod64(w, G.gen_pows[(n - i) % n], q), r32, q);
            }
        }
        T->stepu.alloc(stepu.size() * 4);
        upload_table(ctx, T->stepu, stepu.data(), stepu.size() * 4, "step1 u");
    }
    T->perm.alloc(perm.size() * 4);
    T->w_all.alloc(w_all.size() * 4);
    upload_table(ctx, T->perm, perm.data(), perm.size() * 4, "perm");
    upload_table(ctx, T->w_all, w_all.data(), w_all.size() * 4, "w_all");
    T->gen_pows = G.gen_pows;
    for (uint32_t i = 0; i + 1 < gamma; i++) T->odd_gens &= (G.gen_pows[i] & 1) != 0;
    auto& ref = *T;
    packs[gamma] = std::move(T);
    return ref;
}

PackTables* find_pack_tables(inspire_ctx* ctx, uint32_t gamma) {
    std::lock_guard<std::mutex> lk(root(ctx)->packs_mu);
    auto& packs = root(ctx)->packs;
    auto it = packs.find(gamma);
    return it == packs.end() ? nullptr : it->second.get();
}

PackWsSizes pack_ws_sizes(const inspire_ctx* ctx, const PackTables& T, uint32_t Q) {
    const size_t KN = (size_t)ctx->K * ctx->N;
    const uint32_t nparts = T.fast_front ? ctx->N / 64 : 1;  // the fused front writes one b sum per wave
    PackWsSizes z;
    z.rpre = (size_t)Q * T.gamma * KN;
    z.bpart = (size_t)Q * nparts * T.gamma * ctx->K;
    z.cols = T.fast_front ? 0 : (size_t)Q * 2 * T.gamma * KN;
    return z;
}

bool pack_ybody_in_chain(const inspire_ctx* ctx) { return pack_chain_crt_supported(ctx->dev); }

// Device-only packing of a staged group: first dimension, b_poly, step 1 and
// the chain.  y: NTT-domain (Montgomery) client keys, per query y_off/y_mode/
// y_count (device arrays).  out: [Q][2][K][N].
bool pack_mixed_supported(inspire_ctx* ctx, uint32_t gamma) {
    return pack_chain_crt_supported(ctx->dev) && pack_tables(ctx, gamma).fast_front;
}

uint32_t pack_rows_a_residues(inspire_ctx* ctx, uint32_t gamma) {
    const bool rw = (ctx->params.flags & INSPIRE_FLAG_RESIDUE_WISE_AUTOMORPHISM) != 0;
    if (ctx->K != 2 || rw) return ctx->K;
    return pack_tables(ctx, gamma).fast_front ? 1u : ctx->K;
}

void pack_run(inspire_ctx* ctx, uint32_t gamma, uint32_t Q, const uint32_t* rows, const uint64_t* d_qdig,
              const uint32_t* d_y, const uint64_t* d_yoff, const uint32_t* d_ymode, const uint32_t* d_ycount,
              uint32_t* d_out, bool y_coeff, const PackWs& ws, const PackRoles& roles) {
    const uint32_t K = ctx->K;
    PackTables& T = pack_tables(ctx, gamma);
    if (roles.role && !pack_mixed_supported(ctx, gamma))
        throw Error{INSPIRE_ERR_UNSUPPORTED, "InsPIRe^0 roles in a packing launch need the fused front end"};
    const bool rw = (ctx->params.flags & INSPIRE_FLAG_RESIDUE_WISE_AUTOMORPHISM) != 0;
    const PackWsSizes z = pack_ws_sizes(ctx, T, Q);
    const uint32_t nparts = T.fast_front ? ctx->N / 64 : 1;
    uint32_t* d_rpre = ws.rpre;
    uint64_t* d_bpart = ws.bpart;
    if (!d_rpre || !d_bpart) {
        const size_t rpre_words = z.rpre + (z.rpre & 1);
        ctx->ws3.reserve((rpre_words + 2 * z.bpart) * 4);
        d_rpre = ctx->ws3.get<uint32_t>();
        d_bpart = reinterpret_cast<uint64_t*>(d_rpre + rpre_words);
    }
    if (T.fast_front) {
        const uint32_t* twist = T.twist.get<uint32_t>();
        const uint2* dftw = T.dftw.get<uint2>();
        const bool both = rw || K == 1;
        prof_launch(ctx, P_FRONT, "pack front", [=, dev = ctx->dev, st = ctx->stream](LaunchEv ev) {
            return launch_pack_front(d_qdig, rows, d_rpre, d_bpart, twist, dftw, Q, gamma, both, dev, st, ev, roles,
                                     d_out);
        });
    } else {
        uint32_t* d_cols = ws.cols;
        if (!d_cols) {
            ctx->ws2.reserve(z.cols * 4);
            d_cols = ctx->ws2.get<uint32_t>();
        }
        {
            ProfScope ps(ctx, P_FIRST_DIM);
            hip_check(launch_first_dim(d_qdig, rows, d_cols, Q, gamma, 2 * gamma, 0, ctx->dev, ctx->stream),
                      "first dim");
        }
        {
            ProfScope ps(ctx, P_B0);
            hip_check(launch_pack_b0(d_cols, d_bpart, Q, gamma, ctx->dev, ctx->stream), "b0");
        }
        ProfScope ps(ctx, P_STEP1);
        const uint32_t kres = (rw || K == 1) ? K : 1;  // residue 1 is rebuilt from residue 0 otherwise
        hip_check(launch_pack_step1(d_cols, T.stepu.get<uint32_t>(), d_rpre, Q, gamma, kres,
                                    make_uint2(T.ginvm[0], T.ginvm[1]), ctx->dev, ctx->stream),
                  "step1");
    }
    PackArgs A{};
    A.r_pre = d_rpre;
    A.perm = T.perm.get<uint32_t>();
    A.w_all = T.w_all.get<uint32_t>();
    A.y = d_y;
    A.y_off = d_yoff;
    A.y_mode = d_ymode;
    A.y_count = d_ycount;
    A.y_coeff = y_coeff ? 1u : 0u;
    A.bpart = d_bpart;
    A.nparts = nparts;
    A.out = d_out;
    A.gamma = gamma;
    A.residue_wise = rw ? 1u : 0u;
    A.roles = roles;
    for (uint32_t k = 0; k < K; k++) {
        const uint64_t q = ctx->hm[k].q;
        (k ? A.ninv1 : A.ninv0) = (uint32_t)powmod64(ctx->N, q - 2, q);
        A.ninvm[k] = (uint32_t)mulmod64(powmod64(ctx->N, q - 2, q), (1ull << 32) % q, q);
        A.ninvmp[k] = shoup_pre(A.ninvm[k], (uint32_t)q);
        // gadget powers for the chain's recovered last digit, Montgomery form
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
        A.smul[k] = (uint32_t)r32;
        A.rmul[k] = (uint32_t)(k == 1 && !rw ? powmod64(r32, q - 2, q) : r32);
    }
    const uint64_t q0 = ctx->hm[0].q;
    A.q0r64 = (uint32_t)pow2_64_mod(q0);
    A.q0r64p = shoup_pre(A.q0r64, (uint32_t)q0);
    if (K == 2) {
        const uint64_t q1 = ctx->hm[1].q;
        A.q1r64i = (uint32_t)powmod64(pow2_64_mod(q1), q1 - 2, q1);
        A.q1r64ip = shoup_pre(A.q1r64i, (uint32_t)q1);
        A.one_q1p = shoup_pre(1, (uint32_t)q1);
    }
    prof_launch(
        ctx, P_CHAIN, "pack chain",
        [A, Q, dev = ctx->dev, sc = scale_inv(ctx, nullptr), st = ctx->stream](LaunchEv ev) {
            return launch_pack_chain_scaled(A, Q, dev, sc, st, ev);
        },
        true);
}

static void upload_meta(inspire_ctx* ctx, uint32_t Q, const std::vector<uint64_t>& y_off,
                        const std::vector<uint32_t>& y_mode, const std::vector<uint32_t>& y_count, uint64_t** d_yoff,
                        uint32_t** d_ymode, uint32_t** d_ycount) {
    ctx->ws5.reserve((size_t)Q * 16 + 64);
    *d_yoff = ctx->ws5.get<uint64_t>();
    *d_ymode = reinterpret_cast<uint32_t*>(*d_yoff + Q);
    *d_ycount = *d_ymode + Q;
    hip_check(hipMemcpyAsync(*d_yoff, y_off.data(), Q * 8, hipMemcpyHostToDevice, ctx->stream), "y off");
    hip_check(hipMemcpyAsync(*d_ymode, y_mode.data(), Q * 4, hipMemcpyHostToDevice, ctx->stream), "y mode");
    hip_check(hipMemcpyAsync(*d_ycount, y_count.data(), Q * 4, hipMemcpyHostToDevice, ctx->stream), "y count");
}

void pack_run_device_ybody(inspire_ctx* ctx, uint32_t gamma, uint32_t Q, const uint32_t* rows,
                           const uint64_t* d_qdig, const uint32_t* y_body, uint32_t* d_out, const PackWs& ws,
                           uint32_t* y_ws) {
    const uint32_t ell = ctx->ell;
    const size_t KN = (size_t)ctx->K * ctx->N;
    PackTables& T = pack_tables(ctx, gamma);
    if (!T.odd_gens) throw Error{INSPIRE_ERR_UNSUPPORTED, kEvenGeneratorMsg};
    // y layout: query t at t*ell*K*N, rotated from y_body, gamma-1 rotations (the kernel's default)
    if (pack_ybody_in_chain(ctx)) {  // the chain kernel transforms y_body itself
        pack_run(ctx, gamma, Q, rows, d_qdig, y_body, nullptr, nullptr, nullptr, d_out, true, ws);
        return;
    }
    uint32_t* d_y = y_ws;
    if (!d_y) {
        ctx->ws0.reserve((size_t)Q * ell * KN * 4);
        d_y = ctx->ws0.get<uint32_t>();
    }
    uint64_t mont[kMaxK];
    for (uint32_t k = 0; k < ctx->K; k++) mont[k] = (1ull << 32) % ctx->hm[k].q;
    {
        ProfScope ps(ctx, P_Y_NTT);
        hip_check(launch_ntt_strided(y_body, KN, d_y, KN, (size_t)Q * ell, ctx->dev, scale_fwd(ctx, mont), false,
                                     ctx->stream),
                  "y ntt");
    }
    pack_run(ctx, gamma, Q, rows, d_qdig, d_y, nullptr, nullptr, nullptr, d_out, false, ws);
}

void pack_respond_host(inspire_ctx* ctx, uint32_t gamma, const BatchItem* items, uint32_t Q, const uint32_t* rows,
                       const uint64_t* d_qdig, const uint32_t* d_err, const inspire_query* qs, inspire_response* rs,
                       int32_t* status) {
    const uint32_t K = ctx->K, ell = ctx->ell;
    const size_t KN = (size_t)K * ctx->N;
    PackTables& T = pack_tables(ctx, gamma);
    const uint32_t nrot = gamma - 1;
    const bool odd_gens = T.odd_gens;

    // ---- client keys: y_body (rotated on the device) or explicit y_all
    std::vector<uint64_t> y_off(Q);
    std::vector<uint32_t> y_mode(Q), y_count(Q);
    size_t y_words = 0;
    bool all_body = true;
    for (uint32_t t = 0; t < Q; t++) {
        const inspire_query& q = qs[items[t].idx];
        y_off[t] = y_words;
        if (q.y_all_len > 0) {
            y_mode[t] = 1;
            y_count[t] = std::min(q.y_all_len, nrot);
            y_words += (size_t)std::max(y_count[t], 1u) * ell * KN;
        } else if (odd_gens) {  // y_all[i] = tau_{g^i}(y_body): a permutation of NTT(y_body)
            y_mode[t] = 0;
            y_count[t] = nrot;
            y_words += (size_t)ell * KN;
        } else {  // generate_rotations with even exponents, computed as written
            y_mode[t] = 1;
            y_count[t] = nrot;
            y_words += (size_t)std::max(nrot, 1u) * ell * KN;
        }
        all_body &= y_mode[t] == 0;
    }
    ctx->hkeys.reserve(std::max<size_t>(y_words, 1) * 4);
    uint32_t* hy = ctx->hkeys.get<uint32_t>();
    par_ranges(Q, 4, [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; t++) {
            const inspire_query& q = qs[items[t].idx];
            uint32_t* dst = hy + y_off[t];
            if (q.y_all_len > 0) {
                narrow_u64(dst, q.y_all, (size_t)y_count[t] * ell * KN);
            } else if (y_mode[t] == 0) {
                narrow_u64(dst, q.y_body, (size_t)ell * KN);
            } else {
                for (uint32_t i = 0; i < nrot; i++)
                    for (uint32_t d = 0; d < ell; d++) {
                        std::vector<uint64_t> r = coeff_automorphism(ctx, q.y_body + (size_t)d * KN, T.gen_pows[i]);
                        narrow_u64(dst + ((size_t)i * ell + d) * KN, r.data(), KN);
                    }
            }
        }
    });
    ctx->ws0.reserve(std::max<size_t>(y_words, 1) * 4);
    uint32_t* d_y = ctx->ws0.get<uint32_t>();
    hip_check(hipMemcpyAsync(d_y, hy, y_words * 4, hipMemcpyHostToDevice, ctx->stream), "y");
    // y_body alone: the chain kernel transforms it on the fly
    const bool y_coeff = all_body && pack_chain_crt_supported(ctx->dev);
    if (!y_coeff) {
        uint64_t mont[kMaxK];
        for (uint32_t k = 0; k < K; k++) mont[k] = (1ull << 32) % ctx->hm[k].q;
        ProfScope ps(ctx, P_Y_NTT);
        hip_check(launch_ntt(d_y, y_words / KN, ctx->dev, scale_fwd(ctx, mont), false, ctx->stream), "y ntt");
    }
    uint64_t* d_yoff;
    uint32_t *d_ymode, *d_ycount;
    upload_meta(ctx, Q, y_off, y_mode, y_count, &d_yoff, &d_ymode, &d_ycount);
    const size_t out_words = (size_t)Q * 2 * KN;
    DevBuf& outb = ctx->ws6;
    outb.reserve(out_words * 4);
    pack_run(ctx, gamma, Q, rows, d_qdig, d_y, d_yoff, d_ymode, d_ycount, outb.get<uint32_t>(), y_coeff);
    uint32_t herr = 0;
    const uint32_t* hout = fetch_results(ctx, outb.get<uint32_t>(), out_words, d_err, &herr);
    if (herr) {
        set_last_error("seed expansion exhausted its draw budget");
        for (uint32_t t = 0; t < Q; t++) status[items[t].idx] = INSPIRE_ERR_DEVICE;
        return;
    }
    par_ranges(Q, 2, [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; t++) {
            inspire_response& r = rs[items[t].idx];
            widen_u32(r.ciphertext, hout + t * 2 * KN, 2 * KN);
            r.num_column_ciphertexts = 0;  // respond.rs:382-385
        }
    });
}

}  // namespace inspire
