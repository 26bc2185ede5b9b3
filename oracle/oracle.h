/*
 * CPU restatement of the InsPIRe server respond path of igor53627/inspire-rs
 * (Rust crate `inspire` v0.1.2), written in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py,
 * and nowhere else.  The product (inspire-rs_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Data layout is
 * the reference's `Poly.coeffs` layout (src/math/poly.rs:63-77): a polynomial
 * is `uint64_t[K * d]`, residue-major (all d residues mod moduli[0], then all
 * d residues mod moduli[1]).  An RLWE ciphertext is `a` followed by `b`
 * (`uint64_t[2 * K * d]`).  An RGSW ciphertext is its 2*ell rows in order.
 *
 * Parity status of each part (what pins it) is listed in oracle.c's header.
 */
#ifndef INSPIRE_ORACLE_H
#define INSPIRE_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* InspireParams (src/params.rs:134-185), the fields respond reads. */
typedef struct {
    uint32_t d;            /* ring_dim */
    uint32_t K;            /* crt_moduli.len(): 1 or 2 */
    uint64_t moduli[2];    /* crt_moduli */
    uint64_t q;            /* product of crt_moduli */
    uint64_t p;            /* plaintext modulus */
    uint64_t gadget_base;  /* z */
    uint32_t gadget_len;   /* ell */
} or_params;

/* ---- params.rs ---- */
int      or_params_validate(const or_params *pr);              /* params.rs:348-387 */
uint64_t or_delta(const or_params *pr);                        /* params.rs:296-298 */

/* ---- math/ntt.rs: NttContext ---- */
typedef struct or_ntt or_ntt;
or_ntt  *or_ntt_new(uint32_t n, const uint64_t *moduli, uint32_t K);   /* ntt.rs:116-164 */
void     or_ntt_free(or_ntt *c);
uint64_t or_ntt_psi(const or_ntt *c, uint32_t k);                      /* ntt.rs:525-538 */
void     or_ntt_forward(const or_ntt *c, uint64_t *coeffs);            /* ntt.rs:223-246 */
void     or_ntt_inverse(const or_ntt *c, uint64_t *coeffs);            /* ntt.rs:308-325 */
void     or_ntt_twiddles(const or_ntt *c, uint32_t k, uint64_t *psi, uint64_t *psi_inv);
uint64_t or_mont_mul(const or_ntt *c, uint32_t k, uint64_t a, uint64_t b); /* ntt.rs:468-479 */

/* ---- math/crt.rs, math/poly.rs ---- */
uint64_t or_mod_inverse(uint64_t a, uint64_t m);                        /* crt.rs:10-35 */
uint64_t or_crt_compose_2(uint64_t a0, uint64_t a1, uint64_t q0, uint64_t q1, uint64_t inv); /* crt.rs:41-51 */
uint64_t or_poly_coeff(const or_params *pr, const uint64_t *poly, uint32_t i);   /* poly.rs:359-372 */
void     or_poly_from_coeffs(const or_params *pr, const uint64_t *composite, uint64_t *out); /* poly.rs:127-159 */
void     or_poly_mul_ntt(const or_params *pr, const or_ntt *c, const uint64_t *a,
                         const uint64_t *b, uint64_t *out);             /* poly.rs:505-532 */
void     or_poly_negacyclic_schoolbook(const or_params *pr, const uint64_t *a,
                                       const uint64_t *b, uint64_t *out);

/* ---- rgsw/external_product.rs ---- */
void or_gadget_decompose(const or_params *pr, const uint64_t *poly, uint64_t *digits); /* :16-37 */
void or_gadget_reconstruct(const or_params *pr, const uint64_t *digits, uint64_t *out); /* :43-75 */
/* literal: 4 mul_ntt per digit exactly as external_product.rs:88-150 */
void or_external_product(const or_params *pr, const or_ntt *c, const uint64_t *rlwe,
                         const uint64_t *rgsw, uint64_t *out, int literal);

/* ---- pir/respond.rs: InsPIRe^0 ---- */
/* respond (respond.rs:71-115) on one shard with `ncols` column polys.
 * `cols_out` may be NULL.  ncols == 0 is the empty-shard case (respond.rs:84-90). */
void or_respond(const or_params *pr, const or_ntt *c, const uint64_t *shard_polys, uint32_t ncols,
                const uint64_t *rgsw, uint64_t *ct_out, uint64_t *cols_out, int literal);

/* ---- inspiring/inspiring2.rs ---- */
uint32_t or_pack_generator(uint32_t n, uint32_t gamma);                            /* :114-119 */
void     or_automorph_table(uint32_t n, uint32_t t, uint32_t *table);              /* closed form */
int      or_automorph_table_bruteforce(const or_params *pr, const or_ntt *c, uint32_t t,
                                       uint64_t seed, uint32_t *table);            /* :208-263 */
void     or_apply_automorphism(const or_params *pr, const uint64_t *poly, uint32_t g,
                               uint64_t *out);                                     /* galois.rs:22-55 */
void     or_apply_automorphism_ntt(const or_params *pr, const uint64_t *poly_ntt,
                                   const uint32_t *table, uint64_t *out);          /* :279-296 */
void     or_generate_mask(const or_params *pr, const uint8_t seed[32], uint64_t *w_mask); /* :816-835 */

/* PackParams + OfflinePackingKeys for one (params, gamma, w_seed). */
typedef struct or_pack or_pack;
or_pack *or_pack_new(const or_params *pr, const or_ntt *c, uint32_t gamma,
                     const uint8_t w_seed[32]);           /* :107-182 + :448-506 */
/* residue_wise = 1 applies NTT-domain automorphisms to every residue with the
 * same table (a corrected mode; 0 is the reference's behaviour). */
or_pack *or_pack_new_mode(const or_params *pr, const or_ntt *c, uint32_t gamma,
                          const uint8_t w_seed[32], int residue_wise);
void     or_pack_free(or_pack *pk);
const uint64_t *or_pack_w_all_ntt(const or_pack *pk);     /* [(gamma-1)][ell][K][d] */
/* packing_offline (:926-1041): a_ct [gamma][K][d] coefficient polys ->
 * a_hat [K][d], bold_t [(gamma-1)][ell][K][d] (coefficient form). */
void or_packing_offline(const or_pack *pk, const uint64_t *a_ct, uint64_t *a_hat, uint64_t *bold_t);
/* generate_rotations (:1309-1323) */
void or_generate_rotations(const or_pack *pk, const uint64_t *y_body, uint64_t *y_all);
/* packing_online (:1185-1222): out = RLWE (a_hat, b_poly + INTT(sum)) */
void or_packing_online(const or_pack *pk, const uint64_t *a_hat, const uint64_t *bold_t,
                       const uint64_t *y_all, uint32_t y_all_len, const uint64_t *b_poly,
                       uint64_t *ct_out);

/* respond_inspiring (respond.rs:284-386).  Exactly one of y_all / y_body is
 * used: y_all when y_all_len > 0, else y_body when non-NULL.
 * Return: 0 ok; 1 = empty shard (ct_out = 0, reference returns
 * column_ciphertexts = [0]); -3 = "y_all and y_body are both empty". */
int or_respond_inspiring(const or_params *pr, const or_ntt *c, const or_pack *pk,
                         const uint64_t *shard_polys, uint32_t ncols, const uint64_t *rgsw,
                         const uint64_t *y_body, const uint64_t *y_all, uint32_t y_all_len,
                         uint64_t *ct_out);

/* n respond_inspiring calls (y_body keys) spread over the OpenMP threads. */
void or_respond_inspiring_many(const or_params *pr, const or_ntt *c, const or_pack *pk, uint32_t n,
                               const uint64_t *const *shards, uint32_t ncols, const uint64_t *const *rgsw,
                               const uint64_t *const *y_body, uint64_t *const *ct_out, int *rc);
/* respond_inspiring as the reference runs it, nothing hoisted (per query:
 * NTT contexts, literal external products, PackParams::new with the
 * brute-force table of every odd t < 2d, OfflinePackingKeys::generate); one
 * thread.  Same response as or_respond_inspiring; -4 when a brute-force table
 * differs from the closed form.  work[2] (may be NULL): retries, tables. */
int or_respond_inspiring_literal(const or_params *pr, const uint8_t w_seed[32], const uint64_t *shard_polys,
                                 uint32_t ncols, const uint64_t *rgsw, const uint64_t *y_body, uint64_t *ct_out,
                                 uint64_t *work);
void or_respond_inspiring_literal_many(const or_params *pr, const uint8_t w_seed[32], uint32_t n,
                                       const uint64_t *const *shards, uint32_t ncols, const uint64_t *const *rgsw,
                                       const uint64_t *const *y_body, uint64_t *const *ct_out, int *rc);
int or_num_threads(void);

/* ---- tree packing: inspiring/automorph_pack.rs ---- */
/* sample_extract_coeff0 (rlwe/enc.rs:195-208): composite a [d], b */
void or_sample_extract_coeff0(const or_params *pr, const uint64_t *ct, uint64_t *a_out, uint64_t *b_out);
/* homomorphic_automorph (:44-78); ks = KeySwitchingMatrix rows [ks_rows][2][K][d] */
void or_homomorphic_automorph(const or_params *pr, const or_ntt *c, const uint64_t *ct, uint32_t t,
                              const uint64_t *ks, uint32_t ks_rows, uint64_t *out);
/* pack_lwes (:418-487).  galois_keys [nkeys][ell][2][K][d] (ServerCrs.galois_keys).
 * literal = 1 runs the recursion over all d leaves as written; 0 skips subtrees
 * of zero leaves (identical output).  ct_out: a [K][d] then b [K][bdim]; returns
 * bdim (2d for the reference's two-modulus b, :476-484; d otherwise and with
 * residue_wise = 1, the corrected mode). */
uint32_t or_pack_lwes(const or_params *pr, const or_ntt *c, uint32_t n, const uint64_t *lwe_a,
                      const uint64_t *lwe_b, const uint64_t *galois_keys, uint32_t nkeys, int literal,
                      int residue_wise, uint64_t *ct_out);
/* respond_one_packing (respond.rs:213-268); returns bdim, 0 for an empty shard */
uint32_t or_respond_one_packing(const or_params *pr, const or_ntt *c, const uint64_t *shard_polys,
                                uint32_t ncols, const uint64_t *rgsw, const uint64_t *galois_keys,
                                uint32_t nkeys, int literal, int residue_wise, uint64_t *ct_out);
void or_respond_one_packing_many(const or_params *pr, const or_ntt *c, uint32_t n,
                                 const uint64_t *const *shards, uint32_t ncols, const uint64_t *const *rgsw,
                                 const uint64_t *galois_keys, uint32_t nkeys, uint64_t *const *ct_out);

/* ---- ChaCha20Rng (rand_chacha 0.3) + rand 0.8 gen_range ---- */
typedef struct {
    uint32_t key[8];
    uint64_t counter;      /* next block counter */
    uint32_t buf[64];      /* 4 blocks, as rand_chacha refills */
    uint32_t index;        /* BlockRng index, 64 = empty */
} or_chacha;
void     or_chacha20_block(const uint32_t key[8], uint64_t counter, uint64_t stream, uint32_t out[16]);
void     or_chacha_seed(or_chacha *r, const uint8_t seed[32]);
uint32_t or_chacha_next_u32(or_chacha *r);
uint64_t or_chacha_next_u64(or_chacha *r);
uint64_t or_gen_range(or_chacha *r, uint64_t range);                    /* rand 0.8.5 UniformInt<u64> */
int or_zone_accept(uint64_t v, uint64_t range, unsigned bits, uint64_t *out); /* one draw at width bits */
void or_zone_counts(uint64_t range, unsigned bits, uint64_t *counts);        /* accepted preimages per output */
void     or_poly_from_seed(const or_params *pr, const uint8_t seed[32], uint64_t *out); /* poly.rs:278-281 */
/* SeededRgswCiphertext::expand (rgsw/types.rs:315-318): seeds [2ell][32], b [2ell][K][d] -> rows */
void     or_seeded_rgsw_expand(const or_params *pr, const uint8_t *seeds, const uint64_t *b,
                               uint64_t *rgsw_out);

/* ---- client side (fixtures and decrypt checks only; not on the respond path) ---- */
void or_sample_gaussian_poly(const or_params *pr, or_chacha *r, double sigma, uint64_t *out);
void or_rgsw_encrypt(const or_params *pr, const or_ntt *c, const uint64_t *sk, const uint64_t *msg,
                     or_chacha *r, double sigma, uint64_t *rgsw_out);        /* rgsw/types.rs:133-187 */
void or_seeded_rgsw_encrypt(const or_params *pr, const or_ntt *c, const uint64_t *sk,
                            const uint64_t *msg, or_chacha *r, double sigma,
                            uint8_t *seeds_out, uint64_t *b_out);           /* rgsw/types.rs:235-300 */
void or_inverse_monomial(const or_params *pr, uint32_t k, uint64_t *out);   /* encode_db.rs:80-91 */
void or_client_packing_keys(const or_params *pr, const or_ntt *c, const or_pack *pk,
                            const uint64_t *sk, or_chacha *r, double sigma,
                            uint64_t *y_body, uint64_t *y_all);             /* inspiring2.rs:640-707 */
/* generate_automorphism_ks_matrix (ks/setup.rs:218-247): [ell][2][K][d] */
void or_automorphism_ks_matrix(const or_params *pr, const or_ntt *c, const uint64_t *sk, uint32_t t,
                               or_chacha *r, double sigma, uint64_t *out);
/* ServerCrs.galois_keys (setup.rs:167-173): [log_d][ell][2][K][d] */
void or_galois_keys(const or_params *pr, const or_ntt *c, const uint64_t *sk, or_chacha *r, double sigma,
                    uint64_t *out);
void or_decrypt(const or_params *pr, const or_ntt *c, const uint64_t *ct, const uint64_t *sk,
                uint64_t *msg_out);                                         /* rlwe/enc.rs:90-108 */
/* encode_database (encode_db.rs:106-164) for shard `shard_idx`: writes ncols polys. */
void or_encode_shard(const or_params *pr, const uint8_t *db, uint64_t num_entries, uint32_t entry_size,
                     uint64_t entries_per_shard, uint64_t shard_idx, uint64_t *polys_out);

#ifdef __cplusplus
}
#endif
#endif
