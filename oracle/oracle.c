/*
 * CPU restatement of the InsPIRe respond path of igor53627/inspire-rs.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Plain C, exact integer arithmetic
 * with unsigned __int128 where the reference uses u128.
 *
 * What pins each part (the checks live in tests/test_oracle.py):
 *   NTT (ntt.rs)            closed form a(psi^(2br(k)+1))*2^64 mod q, schoolbook
 *                           negacyclic product, round trip, x*x^(n-1) = -1
 *   CRT / Poly (poly.rs)    closed form; compose(split(x)) = x
 *   gadget (external_product.rs:16-75)  reconstruct identity, digit bounds
 *   external product        schoolbook sum_i D_i*C_i; decrypt invariants
 *   respond (InsPIRe^0)     end-to-end decrypt = entry (d=256 single modulus
 *                           and d=2048 CRT)
 *   automorph tables        brute force of inspiring2.rs:208-263 (small n)
 *   InspiRING, 1 modulus    end-to-end decrypt = entry (the reference's own
 *                           e2e shape, tests/e2e_pir.rs:467-535)
 *   ChaCha20 core           RFC 8439 block and keystream vectors
 * Parity unpinned (no reference output exists to check against; restated from
 * the published source of the dependency / reference):
 *   rand 0.8.5 UniformInt<u64>::sample_single (gen_range) on ChaCha20Rng;
 *   the 2-modulus InspiRING outputs of apply_automorphism_ntt's residue-1
 *   behaviour (inspiring2.rs:279-296), which only a run of the reference pins.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* params.rs                                                            */
/* ------------------------------------------------------------------ */

static uint64_t gcd_u64(uint64_t a, uint64_t b) { /* params.rs:390-397 */
    while (b) { uint64_t t = a % b; a = b; b = t; }
    return a;
}

/* params.rs:348-387 -- returns 0 when valid, a positive code per error string. */
int or_params_validate(const or_params *pr) {
    if (pr->d == 0 || (pr->d & (pr->d - 1))) return 1;      /* "ring_dim must be a power of two" */
    if (pr->K == 0) return 2;                               /* "crt_moduli must be non-empty" */
    if (pr->K > 2) return 3;                                /* "crt_moduli length > 2 is not supported" */
    uint64_t two_n = 2ull * pr->d;
    for (uint32_t k = 0; k < pr->K; k++)
        if (pr->moduli[k] % two_n != 1) return 4;           /* "CRT moduli must be ≡ 1 (mod 2d) for NTT" */
    u128 prod = 1;
    for (uint32_t k = 0; k < pr->K; k++) prod *= pr->moduli[k];
    if ((u128)pr->q != prod) return 5;                      /* "q must equal product of CRT moduli" */
    if (pr->K == 2 && gcd_u64(pr->moduli[0], pr->moduli[1]) != 1) return 6;
    if (pr->q < pr->p) return 7;                            /* "q must be >= p" */
    return 0;
}

uint64_t or_delta(const or_params *pr) { return pr->q / pr->p; } /* params.rs:296-298 */

/* ------------------------------------------------------------------ */
/* math/ntt.rs                                                          */
/* ------------------------------------------------------------------ */

struct or_ntt {
    uint32_t n, K;
    uint64_t q[2], qinv_neg[2], r2[2], n_inv[2], psi[2];
    uint64_t *psi_pow[2], *psi_inv_pow[2];
};

static uint64_t mont_mul_q(uint64_t a, uint64_t b, uint64_t q, uint64_t qinv_neg) { /* ntt.rs:468-479 */
    u128 ab = (u128)a * b;
    uint64_t m = (uint64_t)ab * qinv_neg;
    uint64_t t = (uint64_t)((ab + (u128)m * q) >> 64);
    return t >= q ? t - q : t;
}

uint64_t or_mont_mul(const or_ntt *c, uint32_t k, uint64_t a, uint64_t b) {
    return mont_mul_q(a, b, c->q[k], c->qinv_neg[k]);
}

static uint64_t to_mont(uint64_t a, uint64_t q, uint64_t r2, uint64_t qinv) { /* ntt.rs:481-490 */
    return mont_mul_q(a, r2, q, qinv);
}

static uint64_t compute_q_inv_neg(uint64_t q) { /* ntt.rs:497-504 */
    uint64_t y = 1;
    for (int i = 1; i < 64; i++) y |= (y * q) & (1ull << i);
    return (uint64_t)0 - y;
}

static uint64_t compute_r_squared(uint64_t q) { /* ntt.rs:506-509 */
    u128 r = (((u128)1) << 64) % q;
    return (uint64_t)((r * r) % q);
}

static uint64_t mod_pow(uint64_t base, uint64_t exp, uint64_t m) { /* ntt.rs:511-522 */
    uint64_t result = 1;
    base %= m;
    while (exp) {
        if (exp & 1) result = (uint64_t)(((u128)result * base) % m);
        exp >>= 1;
        base = (uint64_t)(((u128)base * base) % m);
    }
    return result;
}

static uint64_t find_primitive_root(uint64_t n, uint64_t q) { /* ntt.rs:525-538 */
    uint64_t e = (q - 1) / n;
    for (uint64_t g = 2; g < q; g++) {
        uint64_t cand = mod_pow(g, e, q);
        if (mod_pow(cand, n, q) == 1 && mod_pow(cand, n / 2, q) != 1) return cand;
    }
    abort();
}

/* ntt.rs:541-583: factors[m] = psi^bitrev(m) (Montgomery form), built level by level. */
static void compute_twiddles(uint32_t n, uint64_t psi_m, uint64_t q, uint64_t qinv, uint64_t r2,
                             uint64_t *f) {
    f[0] = 0;
    f[1] = to_mont(1, q, r2, qinv);
    for (uint32_t m = 1; m < n; m++) {
        if ((m & (m - 1)) == 0) {
            uint32_t e = n / (2 * m);
            uint64_t pw = to_mont(1, q, r2, qinv);
            for (uint32_t s = 0; s < e; s++) pw = mont_mul_q(pw, psi_m, q, qinv);
            f[m] = pw;
        } else {
            uint32_t prev = m & (m - 1);
            uint32_t step = m & (~m + 1);
            f[m] = mont_mul_q(f[prev], f[step], q, qinv);
        }
    }
}

or_ntt *or_ntt_new(uint32_t n, const uint64_t *moduli, uint32_t K) { /* ntt.rs:116-164 */
    or_ntt *c = (or_ntt *)calloc(1, sizeof(or_ntt));
    c->n = n;
    c->K = K;
    for (uint32_t k = 0; k < K; k++) {
        uint64_t q = moduli[k];
        c->q[k] = q;
        c->qinv_neg[k] = compute_q_inv_neg(q);
        c->r2[k] = compute_r_squared(q);
        uint64_t psi = find_primitive_root(2ull * n, q);
        c->psi[k] = psi;
        c->psi_pow[k] = (uint64_t *)malloc(sizeof(uint64_t) * n);
        c->psi_inv_pow[k] = (uint64_t *)malloc(sizeof(uint64_t) * n);
        compute_twiddles(n, to_mont(psi, q, c->r2[k], c->qinv_neg[k]), q, c->qinv_neg[k], c->r2[k],
                         c->psi_pow[k]);
        uint64_t psi_inv = mod_pow(psi, q - 2, q);
        compute_twiddles(n, to_mont(psi_inv, q, c->r2[k], c->qinv_neg[k]), q, c->qinv_neg[k],
                         c->r2[k], c->psi_inv_pow[k]);
        c->n_inv[k] = to_mont(mod_pow(n, q - 2, q), q, c->r2[k], c->qinv_neg[k]);
    }
    return c;
}

void or_ntt_free(or_ntt *c) {
    if (!c) return;
    for (uint32_t k = 0; k < c->K; k++) { free(c->psi_pow[k]); free(c->psi_inv_pow[k]); }
    free(c);
}

uint64_t or_ntt_psi(const or_ntt *c, uint32_t k) { return c->psi[k]; }

void or_ntt_twiddles(const or_ntt *c, uint32_t k, uint64_t *psi, uint64_t *psi_inv) {
    memcpy(psi, c->psi_pow[k], sizeof(uint64_t) * c->n);
    memcpy(psi_inv, c->psi_inv_pow[k], sizeof(uint64_t) * c->n);
}

static void forward_at(const or_ntt *c, uint64_t *a, uint32_t k) { /* ntt.rs:269-294 */
    const uint32_t n = c->n;
    const uint64_t q = c->q[k], qi = c->qinv_neg[k];
    const uint64_t *w = c->psi_pow[k];
    uint32_t t = n;
    for (uint32_t m = 1; m < n; m <<= 1) {
        t >>= 1;
        for (uint32_t i = 0; i < m; i++) {
            uint32_t j1 = 2 * i * t, j2 = j1 + t;
            uint64_t wi = w[m + i];
            for (uint32_t j = j1; j < j2; j++) {
                uint64_t u = a[j];
                uint64_t v = mont_mul_q(a[j + t], wi, q, qi);
                a[j] = (u + v >= q) ? u + v - q : u + v;
                a[j + t] = (u >= v) ? u - v : q - v + u;
            }
        }
    }
}

static void inverse_at(const or_ntt *c, uint64_t *a, uint32_t k) { /* ntt.rs:348-379 */
    const uint32_t n = c->n;
    const uint64_t q = c->q[k], qi = c->qinv_neg[k];
    const uint64_t *w = c->psi_inv_pow[k];
    uint32_t t = 1;
    for (uint32_t m = n; m > 1;) {
        m >>= 1;
        for (uint32_t i = 0; i < m; i++) {
            uint32_t j2 = i * 2 * t;
            uint64_t wi = w[m + i];
            for (uint32_t j = j2; j < j2 + t; j++) {
                uint64_t u = a[j], v = a[j + t];
                a[j] = (u + v >= q) ? u + v - q : u + v;
                uint64_t diff = (u >= v) ? u - v : q - v + u;
                a[j + t] = mont_mul_q(diff, wi, q, qi);
            }
        }
        t <<= 1;
    }
    for (uint32_t j = 0; j < n; j++) a[j] = mont_mul_q(a[j], c->n_inv[k], q, qi);
}

void or_ntt_forward(const or_ntt *c, uint64_t *coeffs) { /* ntt.rs:223-246 */
    for (uint32_t k = 0; k < c->K; k++) {
        uint64_t *a = coeffs + (size_t)k * c->n;
        for (uint32_t j = 0; j < c->n; j++) a[j] = to_mont(a[j], c->q[k], c->r2[k], c->qinv_neg[k]);
        forward_at(c, a, k);
    }
}

void or_ntt_inverse(const or_ntt *c, uint64_t *coeffs) { /* ntt.rs:308-325 */
    for (uint32_t k = 0; k < c->K; k++) inverse_at(c, coeffs + (size_t)k * c->n, k);
    for (uint32_t k = 0; k < c->K; k++) {
        uint64_t *a = coeffs + (size_t)k * c->n;
        for (uint32_t j = 0; j < c->n; j++) a[j] = mont_mul_q(a[j], 1, c->q[k], c->qinv_neg[k]);
    }
}

/* ------------------------------------------------------------------ */
/* math/crt.rs, math/poly.rs                                            */
/* ------------------------------------------------------------------ */

uint64_t or_mod_inverse(uint64_t a, uint64_t modulus) { /* crt.rs:10-35 */
    __int128 t = 0, new_t = 1, r = modulus, new_r = a;
    while (new_r != 0) {
        __int128 quo = r / new_r;
        __int128 tmp_t = t - quo * new_t; t = new_t; new_t = tmp_t;
        __int128 tmp_r = r - quo * new_r; r = new_r; new_r = tmp_r;
    }
    if (r != 1) abort();
    if (t < 0) t += modulus;
    return (uint64_t)t;
}

uint64_t or_crt_compose_2(uint64_t a0, uint64_t a1, uint64_t q0, uint64_t q1, uint64_t inv) {
    uint64_t a0m = a0 % q1;                                      /* crt.rs:41-51 */
    uint64_t diff = (a1 >= a0m) ? a1 - a0m : (a1 + q1) - a0m;
    uint64_t t = (uint64_t)(((u128)diff * inv) % q1);
    return a0 + q0 * t;
}

static uint64_t crt_inv(const or_params *pr) {
    return pr->K == 2 ? or_mod_inverse(pr->moduli[0], pr->moduli[1]) : 0; /* poly.rs:80-93 */
}

uint64_t or_poly_coeff(const or_params *pr, const uint64_t *poly, uint32_t i) { /* poly.rs:359-372 */
    if (pr->K == 1) return poly[i];
    return or_crt_compose_2(poly[i], poly[i + pr->d], pr->moduli[0], pr->moduli[1], crt_inv(pr)) % pr->q;
}

static uint64_t poly_coeff_inv(const or_params *pr, const uint64_t *poly, uint32_t i, uint64_t inv) {
    if (pr->K == 1) return poly[i];
    return or_crt_compose_2(poly[i], poly[i + pr->d], pr->moduli[0], pr->moduli[1], inv) % pr->q;
}

static void poly_set_coeff(const or_params *pr, uint64_t *poly, uint32_t i, uint64_t v) { /* poly.rs:375-388 */
    for (uint32_t k = 0; k < pr->K; k++) poly[i + (size_t)k * pr->d] = v % pr->moduli[k];
}

void or_poly_from_coeffs(const or_params *pr, const uint64_t *composite, uint64_t *out) { /* poly.rs:127-159 */
    for (uint32_t k = 0; k < pr->K; k++)
        for (uint32_t i = 0; i < pr->d; i++) out[(size_t)k * pr->d + i] = composite[i] % pr->moduli[k];
}

static void poly_add(const or_params *pr, const uint64_t *a, const uint64_t *b, uint64_t *out) { /* poly.rs:710-736 */
    for (uint32_t k = 0; k < pr->K; k++) {
        uint64_t q = pr->moduli[k];
        for (uint32_t i = 0; i < pr->d; i++) {
            size_t x = (size_t)k * pr->d + i;
            uint64_t s = a[x] + b[x];
            out[x] = s >= q ? s - q : s;
        }
    }
}

static void poly_neg(const or_params *pr, const uint64_t *a, uint64_t *out) { /* poly.rs:807-829 */
    for (uint32_t k = 0; k < pr->K; k++) {
        uint64_t q = pr->moduli[k];
        for (uint32_t i = 0; i < pr->d; i++) {
            size_t x = (size_t)k * pr->d + i;
            out[x] = a[x] == 0 ? 0 : q - a[x];
        }
    }
}

static void poly_scalar_mul(const or_params *pr, const uint64_t *a, uint64_t s, uint64_t *out) { /* poly.rs:466-485 */
    for (uint32_t k = 0; k < pr->K; k++) {
        uint64_t q = pr->moduli[k], sm = s % q;
        for (uint32_t i = 0; i < pr->d; i++) {
            size_t x = (size_t)k * pr->d + i;
            out[x] = (uint64_t)(((u128)a[x] * sm) % q);
        }
    }
}

/* pointwise_mul (ntt.rs:394-417) / mul_ntt_domain (poly.rs:535-553) */
static void poly_pointwise(const or_params *pr, const or_ntt *c, const uint64_t *a, const uint64_t *b,
                           uint64_t *out) {
    for (uint32_t k = 0; k < pr->K; k++)
        for (uint32_t i = 0; i < pr->d; i++) {
            size_t x = (size_t)k * pr->d + i;
            out[x] = mont_mul_q(a[x], b[x], c->q[k], c->qinv_neg[k]);
        }
}

/* mul_acc_ntt_domain (poly.rs:613-630): acc += a*b */
static void poly_mul_acc(const or_params *pr, const or_ntt *c, uint64_t *acc, const uint64_t *a,
                         const uint64_t *b) {
    for (uint32_t k = 0; k < pr->K; k++) {
        uint64_t q = c->q[k];
        for (uint32_t i = 0; i < pr->d; i++) {
            size_t x = (size_t)k * pr->d + i;
            uint64_t s = acc[x] + mont_mul_q(a[x], b[x], q, c->qinv_neg[k]);
            acc[x] = s >= q ? s - q : s;
        }
    }
}

static size_t psz(const or_params *pr) { return (size_t)pr->K * pr->d; }
static uint64_t *palloc(const or_params *pr, size_t npolys) {
    return (uint64_t *)calloc(npolys * psz(pr), sizeof(uint64_t));
}

void or_poly_mul_ntt(const or_params *pr, const or_ntt *c, const uint64_t *a, const uint64_t *b,
                     uint64_t *out) { /* poly.rs:505-532 */
    uint64_t *ta = palloc(pr, 1), *tb = palloc(pr, 1);
    memcpy(ta, a, psz(pr) * 8);
    memcpy(tb, b, psz(pr) * 8);
    or_ntt_forward(c, ta);
    or_ntt_forward(c, tb);
    poly_pointwise(pr, c, ta, tb, out);
    or_ntt_inverse(c, out);
    free(ta);
    free(tb);
}

/* Brute-force negacyclic product, residue by residue: an independent check of the NTT. */
void or_poly_negacyclic_schoolbook(const or_params *pr, const uint64_t *a, const uint64_t *b,
                                   uint64_t *out) {
    const uint32_t n = pr->d;
    for (uint32_t k = 0; k < pr->K; k++) {
        const uint64_t q = pr->moduli[k];
        const uint64_t *x = a + (size_t)k * n, *y = b + (size_t)k * n;
        uint64_t *z = out + (size_t)k * n;
        for (uint32_t i = 0; i < n; i++) {
            u128 pos = 0, neg = 0;
            for (uint32_t j = 0; j <= i; j++) pos += ((u128)x[j] * y[i - j]) % q;
            for (uint32_t j = i + 1; j < n; j++) neg += ((u128)x[j] * y[n + i - j]) % q;
            uint64_t p = (uint64_t)(pos % q), m = (uint64_t)(neg % q);
            z[i] = p >= m ? p - m : q - m + p;
        }
    }
}

/* ------------------------------------------------------------------ */
/* rgsw: gadget decomposition and the external product                 */
/* ------------------------------------------------------------------ */

void or_gadget_decompose(const or_params *pr, const uint64_t *poly, uint64_t *digits) {
    /* external_product.rs:16-37: digits are [ell][K][d] */
    const uint64_t inv = crt_inv(pr);
    memset(digits, 0, sizeof(uint64_t) * pr->gadget_len * psz(pr));
    for (uint32_t j = 0; j < pr->d; j++) {
        uint64_t val = poly_coeff_inv(pr, poly, j, inv);
        for (uint32_t i = 0; i < pr->gadget_len; i++) {
            poly_set_coeff(pr, digits + (size_t)i * psz(pr), j, val % pr->gadget_base);
            val /= pr->gadget_base;
        }
    }
}

static void gadget_powers(const or_params *pr, uint64_t *pw) { /* rgsw/types.rs:79-90 */
    u128 cur = 1;
    for (uint32_t i = 0; i < pr->gadget_len; i++) {
        pw[i] = (uint64_t)cur;
        cur = (cur * pr->gadget_base) % pr->q;
    }
}

void or_gadget_reconstruct(const or_params *pr, const uint64_t *digits, uint64_t *out) { /* :43-75 */
    uint64_t pw[16];
    gadget_powers(pr, pw);
    uint64_t *t = palloc(pr, 1);
    memset(out, 0, psz(pr) * 8);
    for (uint32_t i = 0; i < pr->gadget_len; i++) {
        poly_scalar_mul(pr, digits + (size_t)i * psz(pr), pw[i], t);
        poly_add(pr, out, t, out);
    }
    free(t);
}

/* external_product.rs:88-150.  rgsw rows: [2ell][2][K][d]; rlwe/out: [2][K][d]. */
void or_external_product(const or_params *pr, const or_ntt *c, const uint64_t *rlwe,
                         const uint64_t *rgsw, uint64_t *out, int literal) {
    const size_t P = psz(pr);
    const uint32_t ell = pr->gadget_len;
    uint64_t *ad = palloc(pr, ell), *bd = palloc(pr, ell);
    or_gadget_decompose(pr, rlwe, ad);
    or_gadget_decompose(pr, rlwe + P, bd);
    uint64_t *ra = out, *rb = out + P;
    memset(out, 0, 2 * P * 8);
    if (literal) {
        uint64_t *t = palloc(pr, 1);
        for (uint32_t i = 0; i < ell; i++) {
            const uint64_t *row_a = rgsw + (size_t)(2 * i) * P;          /* rows[i] */
            const uint64_t *row_b = rgsw + (size_t)(2 * (ell + i)) * P;  /* rows[ell+i] */
            or_poly_mul_ntt(pr, c, ad + i * P, row_a, t);     poly_add(pr, ra, t, ra);
            or_poly_mul_ntt(pr, c, ad + i * P, row_a + P, t); poly_add(pr, rb, t, rb);
            or_poly_mul_ntt(pr, c, bd + i * P, row_b, t);     poly_add(pr, ra, t, ra);
            or_poly_mul_ntt(pr, c, bd + i * P, row_b + P, t); poly_add(pr, rb, t, rb);
        }
        free(t);
    } else {
        /* Same sums with each operand transformed once and one inverse per output:
         * INTT is linear and every value is canonical, so the result is identical. */
        uint64_t *dn = palloc(pr, 1), *cn = palloc(pr, 1);
        for (uint32_t i = 0; i < ell; i++) {
            for (int half = 0; half < 2; half++) {
                const uint64_t *dig = (half ? bd : ad) + i * P;
                int zero = 1;
                for (size_t x = 0; x < P && zero; x++) zero = dig[x] == 0;
                if (zero) continue; /* a product with an all-zero digit is exactly 0 */
                memcpy(dn, dig, P * 8);
                or_ntt_forward(c, dn);
                const uint64_t *row = rgsw + (size_t)(2 * (half ? ell + i : i)) * P;
                for (int comp = 0; comp < 2; comp++) {
                    memcpy(cn, row + comp * P, P * 8);
                    or_ntt_forward(c, cn);
                    poly_mul_acc(pr, c, comp ? rb : ra, dn, cn);
                }
            }
        }
        or_ntt_inverse(c, ra);
        or_ntt_inverse(c, rb);
        free(dn);
        free(cn);
    }
    free(ad);
    free(bd);
}

/* ------------------------------------------------------------------ */
/* pir/respond.rs: InsPIRe^0                                            */
/* ------------------------------------------------------------------ */

/* RlweCiphertext::trivial_encrypt (rlwe/enc.rs:174-182) */
static void trivial_encrypt(const or_params *pr, const uint64_t *h, uint64_t *ct) {
    memset(ct, 0, psz(pr) * 8);
    poly_scalar_mul(pr, h, or_delta(pr), ct + psz(pr));
}

void or_respond(const or_params *pr, const or_ntt *c, const uint64_t *shard_polys, uint32_t ncols,
                const uint64_t *rgsw, uint64_t *ct_out, uint64_t *cols_out, int literal) {
    const size_t P = psz(pr);
    if (ncols == 0) { /* respond.rs:84-90 */
        memset(ct_out, 0, 2 * P * 8);
        return;
    }
    uint64_t *cols = cols_out ? cols_out : palloc(pr, 2 * (size_t)ncols);
#pragma omp parallel for schedule(dynamic, 1) if (!literal)
    for (uint32_t col = 0; col < ncols; col++) { /* respond.rs:92-100 */
        uint64_t *rl = palloc(pr, 2);
        trivial_encrypt(pr, shard_polys + (size_t)col * P, rl);
        or_external_product(pr, c, rl, rgsw, cols + (size_t)col * 2 * P, literal);
        free(rl);
    }
    memcpy(ct_out, cols, 2 * P * 8); /* respond.rs:102-109 */
    for (uint32_t col = 1; col < ncols; col++) {
        poly_add(pr, ct_out, cols + (size_t)col * 2 * P, ct_out);
        poly_add(pr, ct_out + P, cols + (size_t)col * 2 * P + P, ct_out + P);
    }
    if (!cols_out) free(cols);
}

/* ------------------------------------------------------------------ */
/* ChaCha20Rng (rand_chacha 0.3.1, rand_core 0.6.4) and rand 0.8.5      */
/* ------------------------------------------------------------------ */

#define ROTL32(v, n) (((v) << (n)) | ((v) >> (32 - (n))))
#define QR(a, b, c, d)                                   \
    a += b; d ^= a; d = ROTL32(d, 16);                   \
    c += d; b ^= c; b = ROTL32(b, 12);                   \
    a += b; d ^= a; d = ROTL32(d, 8);                    \
    c += d; b ^= c; b = ROTL32(b, 7);

/* ChaCha20 block with a 64-bit block counter (words 12-13) and a 64-bit
 * stream id (words 14-15), the layout rand_chacha uses. */
void or_chacha20_block(const uint32_t key[8], uint64_t counter, uint64_t stream, uint32_t out[16]) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; i++) s[4 + i] = key[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = (uint32_t)stream;
    s[15] = (uint32_t)(stream >> 32);
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    for (int r = 0; r < 10; r++) {
        QR(x[0], x[4], x[8], x[12]) QR(x[1], x[5], x[9], x[13])
        QR(x[2], x[6], x[10], x[14]) QR(x[3], x[7], x[11], x[15])
        QR(x[0], x[5], x[10], x[15]) QR(x[1], x[6], x[11], x[12])
        QR(x[2], x[7], x[8], x[13]) QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

void or_chacha_seed(or_chacha *r, const uint8_t seed[32]) { /* ChaCha20Rng::from_seed */
    for (int i = 0; i < 8; i++)
        r->key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                    ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    r->counter = 0;
    r->index = 64;
}

static void chacha_refill(or_chacha *r) { /* 4 blocks per refill, in counter order */
    for (int b = 0; b < 4; b++) or_chacha20_block(r->key, r->counter + b, 0, r->buf + 16 * b);
    r->counter += 4;
}

uint32_t or_chacha_next_u32(or_chacha *r) { /* rand_core BlockRng::next_u32 */
    if (r->index >= 64) { chacha_refill(r); r->index = 0; }
    return r->buf[r->index++];
}

uint64_t or_chacha_next_u64(or_chacha *r) { /* rand_core 0.6 BlockRng::next_u64 */
    uint32_t len = 64, idx = r->index;
    if (idx < len - 1) {
        r->index += 2;
        return (uint64_t)r->buf[idx] | ((uint64_t)r->buf[idx + 1] << 32);
    } else if (idx >= len) {
        chacha_refill(r);
        r->index = 2;
        return (uint64_t)r->buf[0] | ((uint64_t)r->buf[1] << 32);
    } else {
        uint64_t x = r->buf[len - 1];
        chacha_refill(r);
        r->index = 1;
        uint64_t y = r->buf[0];
        return (y << 32) | x;
    }
}

/* One draw of rand 0.8.5's UniformInt sample_single_inclusive for the
 * integer width `bits` (the u64 instance has bits = 64): v uniform in
 * [0, 2^bits), 1 <= range < 2^bits.  The widening product v * range splits
 * into hi (the candidate output) and lo (the low bits); the draw is accepted
 * when lo <= zone, the "conservative zone" (range << leading_zeros(range)) - 1
 * at that width.  Returns 1 and sets *out = hi when accepted.  Written for any
 * width so that the same code can be checked exhaustively at 8 and 16 bits. */
int or_zone_accept(uint64_t v, uint64_t range, unsigned bits, uint64_t *out) {
    const u128 mask = ((u128)1 << bits) - 1;
    const unsigned lz = (unsigned)__builtin_clzll(range) - (64u - bits);
    const uint64_t zone = (uint64_t)((((u128)range << lz) - 1) & mask);
    const u128 m = (u128)v * range;
    const uint64_t hi = (uint64_t)(m >> bits), lo = (uint64_t)(m & mask);
    if (lo <= zone) {
        *out = hi;
        return 1;
    }
    return 0;
}

/* Test support: for every v in [0, 2^bits) (bits <= 24), the accepted
 * outputs of or_zone_accept counted per value: counts[range]. */
void or_zone_counts(uint64_t range, unsigned bits, uint64_t *counts) {
    for (uint64_t i = 0; i < range; i++) counts[i] = 0;
    for (uint64_t v = 0; v < (1ull << bits); v++) {
        uint64_t out;
        if (or_zone_accept(v, range, bits, &out)) counts[out]++;
    }
}

/* rand 0.8.5 UniformInt<u64>::sample_single(0, range) -> sample_single_inclusive(0, range-1):
 * u64 draws until one is accepted. */
uint64_t or_gen_range(or_chacha *r, uint64_t range) {
    uint64_t out;
    for (;;)
        if (or_zone_accept(or_chacha_next_u64(r), range, 64, &out)) return out;
}

/* Poly::random_with_rng_moduli (poly.rs:251-268): residue-major draws. */
static void poly_random_rng(const or_params *pr, or_chacha *r, uint64_t *out) {
    for (uint32_t k = 0; k < pr->K; k++)
        for (uint32_t i = 0; i < pr->d; i++) out[(size_t)k * pr->d + i] = or_gen_range(r, pr->moduli[k]);
}

void or_poly_from_seed(const or_params *pr, const uint8_t seed[32], uint64_t *out) { /* poly.rs:278-281 */
    or_chacha r;
    or_chacha_seed(&r, seed);
    poly_random_rng(pr, &r, out);
}

void or_seeded_rgsw_expand(const or_params *pr, const uint8_t *seeds, const uint64_t *b,
                           uint64_t *rgsw_out) { /* rgsw/types.rs:315-318, rlwe/types.rs:193-197 */
    const size_t P = psz(pr);
    for (uint32_t row = 0; row < 2 * pr->gadget_len; row++) {
        or_poly_from_seed(pr, seeds + 32 * row, rgsw_out + (size_t)(2 * row) * P);
        memcpy(rgsw_out + (size_t)(2 * row + 1) * P, b + (size_t)row * P, P * 8);
    }
}

/* ------------------------------------------------------------------ */
/* inspiring/inspiring2.rs: PackParams, keys, offline/online packing    */
/* ------------------------------------------------------------------ */

uint32_t or_pack_generator(uint32_t n, uint32_t gamma) { /* inspiring2.rs:114-119 */
    return gamma < n ? (2 * n / gamma) + 1 : 5;
}

static uint32_t bitrev(uint32_t x, uint32_t logn) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < logn; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* Closed form of the table generate_automorph_tables (inspiring2.rs:208-263) finds:
 * NTT slot j evaluates at psi^(2 br(j) + 1), and tau_t maps that point to
 * psi^(t (2 br(j) + 1)), so out[j] = in[br(((t (2 br(j)+1)) mod 2n - 1) / 2)]. */
void or_automorph_table(uint32_t n, uint32_t t, uint32_t *table) {
    uint32_t logn = 0;
    while ((1u << logn) < n) logn++;
    for (uint32_t j = 0; j < n; j++) {
        uint64_t e = ((uint64_t)t * (2 * bitrev(j, logn) + 1)) % (2ull * n);
        table[j] = bitrev((uint32_t)((e - 1) / 2), logn);
    }
}

/* galois.rs:22-55, on composite values exactly as written. */
void or_apply_automorphism(const or_params *pr, const uint64_t *poly, uint32_t g, uint64_t *out) {
    const uint32_t d = pr->d;
    const uint64_t q = pr->q, inv = crt_inv(pr);
    uint64_t *res = (uint64_t *)calloc(d, sizeof(uint64_t));
    for (uint32_t i = 0; i < d; i++) {
        uint64_t cf = poly_coeff_inv(pr, poly, i, inv);
        if (cf == 0) continue;
        uint64_t ni = ((uint64_t)g * i) % (2ull * d);
        if (ni < d) {
            res[ni] = (uint64_t)(((u128)res[ni] + cf) % q);          /* mod_add, galois.rs:128-131 */
        } else {
            uint64_t ai = ni - d;                                    /* mod_sub, galois.rs:135-141 */
            res[ai] = res[ai] >= cf ? res[ai] - cf : q - cf + res[ai];
        }
    }
    or_poly_from_coeffs(pr, res, out);
    free(res);
}

/* generate_automorph_tables (inspiring2.rs:208-263) for one t, by brute force
 * over a seeded random polynomial (the reference draws thread_rng; the table it
 * finds is the same unique permutation).  Returns the number of retries. */
int or_automorph_table_bruteforce(const or_params *pr, const or_ntt *c, uint32_t t, uint64_t seed,
                                  uint32_t *table) {
    const uint32_t n = pr->d;
    uint8_t sd[32] = {0};
    uint64_t *p = palloc(pr, 1), *pa = palloc(pr, 1);
    for (int attempt = 0;; attempt++) {
        uint64_t s = seed + (uint64_t)attempt * 0x9E3779B97F4A7C15ull;
        memcpy(sd, &s, 8);
        or_poly_from_seed(pr, sd, p);
        or_apply_automorphism(pr, p, t, pa);
        or_ntt_forward(c, p);
        or_ntt_forward(c, pa);
        int redo = 0;
        for (uint32_t i = 0; i < n && !redo; i++) {
            uint32_t count = 0, found = 0;
            for (uint32_t j = 0; j < n; j++)
                if (pa[j] == p[i]) { count++; found = j; }
            if (count != 1) redo = 1;
            else table[found] = i;
        }
        if (!redo) {
            free(p);
            free(pa);
            return attempt;
        }
    }
}

/* apply_automorphism_ntt (inspiring2.rs:279-296): permute the first n values
 * (residue 0), then Poly::from_coeffs_moduli splits each value v into
 * (v mod q0, v mod q1) -- with two moduli residue 1 is rebuilt from residue 0. */
void or_apply_automorphism_ntt(const or_params *pr, const uint64_t *poly_ntt, const uint32_t *table,
                               uint64_t *out) {
    const uint32_t n = pr->d;
    uint64_t *res = (uint64_t *)malloc(sizeof(uint64_t) * n);
    for (uint32_t i = 0; i < n; i++) res[i] = poly_ntt[table[i]];
    or_poly_from_coeffs(pr, res, out);
    free(res);
}

/* Residue-wise NTT-domain automorphism: the same permutation applied to every
 * residue (the table does not depend on psi).  Not what the reference does
 * with two moduli; offered as a non-default corrected mode. */
static void apply_automorphism_ntt_rw(const or_params *pr, const uint64_t *poly_ntt,
                                      const uint32_t *table, uint64_t *out) {
    for (uint32_t k = 0; k < pr->K; k++)
        for (uint32_t i = 0; i < pr->d; i++)
            out[(size_t)k * pr->d + i] = poly_ntt[(size_t)k * pr->d + table[i]];
}

static void pack_automorph_ntt(const or_pack *pk, const uint64_t *in, const uint32_t *table,
                               uint64_t *out);

/* generate_mask_from_seed (inspiring2.rs:816-835): [ell][K][d] */
void or_generate_mask(const or_params *pr, const uint8_t seed[32], uint64_t *w_mask) {
    or_chacha r;
    or_chacha_seed(&r, seed);
    uint64_t *cf = (uint64_t *)malloc(sizeof(uint64_t) * pr->d);
    for (uint32_t k = 0; k < pr->gadget_len; k++) {
        for (uint32_t i = 0; i < pr->d; i++) cf[i] = or_chacha_next_u64(&r) % pr->q;
        or_poly_from_coeffs(pr, cf, w_mask + (size_t)k * psz(pr));
    }
    free(cf);
}

struct or_pack {
    or_params pr;
    int residue_wise;       /* 0: the reference's apply_automorphism_ntt; 1: corrected */
    const or_ntt *c;
    uint32_t gamma, generator;
    uint32_t *gen_pows;     /* [n] */
    uint64_t mod_inv_gamma;
    uint64_t *mod_inv_ntt;  /* [K][d] */
    uint32_t *tables;       /* [gamma][n]: table for gen_pows[i] */
    uint64_t *mono;         /* [gamma][gamma][K][d]: NTT(X^e_ij), e_ij = j g^((n-i) mod n) mod 2n */
    uint64_t *w_mask;       /* [ell][K][d], coefficient form */
    uint64_t *w_all_ntt;    /* [(gamma-1)][ell][K][d] */
};

static uint64_t mod_inverse_u64(uint64_t a, uint64_t m) { /* inspiring2.rs:1633-1640 */
    __int128 old_r = a, r = m, old_s = 1, s = 0;
    while (r != 0) {
        __int128 qq = old_r / r, t;
        t = r; r = old_r - qq * r; old_r = t;
        t = s; s = old_s - qq * s; old_s = t;
    }
    if (old_r != 1) abort();
    __int128 x = old_s % (__int128)m;
    if (x < 0) x += m;
    return (uint64_t)x;
}

static uint32_t table_index(uint32_t t) { return (t - 1) / 2; } /* inspiring2.rs:188-192 */

static void pack_automorph_ntt(const or_pack *pk, const uint64_t *in, const uint32_t *table,
                               uint64_t *out) {
    if (pk->residue_wise) apply_automorphism_ntt_rw(&pk->pr, in, table, out);
    else or_apply_automorphism_ntt(&pk->pr, in, table, out);
}

/* NTT form of the monomial X^e, e < 2n, built as PackParams::new does
 * (inspiring2.rs:143-158): from_coeffs_moduli then to_ntt.  X^e for e >= n is
 * the negated monomial (coefficient q - 1 at e - n). */
static void monomial_ntt(const or_pack *pk, uint32_t e, uint64_t *out) {
    const or_params *pr = &pk->pr;
    const uint32_t n = pr->d;
    uint64_t *cf = (uint64_t *)calloc(n, sizeof(uint64_t));
    cf[e % n] = e < n ? 1 : pr->q - 1;
    or_poly_from_coeffs(pr, cf, out);
    or_ntt_forward(pk->c, out);
    free(cf);
}

or_pack *or_pack_new(const or_params *pr, const or_ntt *c, uint32_t gamma, const uint8_t w_seed[32]) {
    return or_pack_new_mode(pr, c, gamma, w_seed, 0);
}

/* PackParams::new's scalars (inspiring2.rs:115-136): generator, its n powers
 * mod 2n, mod_inv_gamma and its NTT as a constant polynomial. */
static or_pack *pack_alloc(const or_params *pr, const or_ntt *c, uint32_t gamma, int residue_wise) {
    const uint32_t n = pr->d, two_n = 2 * n;
    or_pack *pk = (or_pack *)calloc(1, sizeof(or_pack));
    pk->pr = *pr;
    pk->residue_wise = residue_wise;
    pk->c = c;
    pk->gamma = gamma;
    pk->generator = or_pack_generator(n, gamma);
    pk->gen_pows = (uint32_t *)malloc(sizeof(uint32_t) * n);      /* inspiring2.rs:122-128 */
    uint64_t v = 1;
    for (uint32_t i = 0; i < n; i++) { pk->gen_pows[i] = (uint32_t)v; v = (v * pk->generator) % two_n; }
    pk->mod_inv_gamma = mod_inverse_u64(gamma, pr->q);             /* :131-132 */
    pk->mod_inv_ntt = palloc(pr, 1);                               /* :135-136 */
    uint64_t *cf = (uint64_t *)calloc(n, sizeof(uint64_t));
    cf[0] = pk->mod_inv_gamma;
    or_poly_from_coeffs(pr, cf, pk->mod_inv_ntt);
    or_ntt_forward(c, pk->mod_inv_ntt);
    free(cf);
    pk->tables = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)gamma * n);
    pk->mono = palloc(pr, (size_t)gamma * gamma);
    return pk;
}

/* OfflinePackingKeys::generate (inspiring2.rs:448-506): the mask from the
 * seed, its NTT, and its rotations by the tables of g^i, i < gamma - 1. */
static void pack_keys(or_pack *pk, const uint8_t w_seed[32]) {
    const or_params *pr = &pk->pr;
    const uint32_t n = pr->d, ell = pr->gadget_len, gamma = pk->gamma;
    const size_t P = psz(pr);
    pk->w_mask = palloc(pr, ell);
    or_generate_mask(pr, w_seed, pk->w_mask);
    uint64_t *w_mask = palloc(pr, ell);
    memcpy(w_mask, pk->w_mask, (size_t)ell * P * 8);
    for (uint32_t k = 0; k < ell; k++) or_ntt_forward(pk->c, w_mask + (size_t)k * P);
    uint32_t nrot = gamma > 0 ? gamma - 1 : 0;
    pk->w_all_ntt = palloc(pr, (size_t)(nrot ? nrot : 1) * ell);
    for (uint32_t i = 0; i < nrot; i++)
        for (uint32_t k = 0; k < ell; k++)
            pack_automorph_ntt(pk, w_mask + (size_t)k * P, pk->tables + (size_t)i * n,
                               pk->w_all_ntt + ((size_t)i * ell + k) * P);
    free(w_mask);
}

/* The query-independent tables built once (the hoisted form): the closed-form
 * automorphism tables of the gamma generator powers packing reads, and the
 * gamma^2 NTT-form monomials step 1 reads (PackParams.monomials_ntt,
 * inspiring2.rs:143-158), kept per (i, j). */
or_pack *or_pack_new_mode(const or_params *pr, const or_ntt *c, uint32_t gamma, const uint8_t w_seed[32],
                          int residue_wise) {
    const uint32_t n = pr->d;
    const size_t P = psz(pr);
    or_pack *pk = pack_alloc(pr, c, gamma, residue_wise);
    for (uint32_t i = 0; i < gamma; i++)
        or_automorph_table(n, 2 * table_index(pk->gen_pows[i]) + 1, pk->tables + (size_t)i * n);
    for (uint32_t i = 0; i < gamma; i++)
        for (uint32_t j = 0; j < gamma; j++) {
            uint32_t e = (uint32_t)(((uint64_t)j * pk->gen_pows[(n - i) % n]) % (2ull * n));
            monomial_ntt(pk, e, pk->mono + ((size_t)i * gamma + j) * P);
        }
    pack_keys(pk, w_seed);
    return pk;
}

void or_pack_free(or_pack *pk) {
    if (!pk) return;
    free(pk->gen_pows);
    free(pk->mod_inv_ntt);
    free(pk->tables);
    free(pk->mono);
    free(pk->w_all_ntt);
    free(pk->w_mask);
    free(pk);
}

const uint64_t *or_pack_w_all_ntt(const or_pack *pk) { return pk->w_all_ntt; }

void or_packing_offline(const or_pack *pk, const uint64_t *a_ct, uint64_t *a_hat, uint64_t *bold_t) {
    const or_params *pr = &pk->pr;
    const or_ntt *c = pk->c;
    const uint32_t n = pr->d, gamma = pk->gamma, ell = pr->gadget_len;
    const size_t P = psz(pr);
    uint64_t *a_ntt = palloc(pr, gamma), *r_all = palloc(pr, gamma);
    uint64_t *mono = palloc(pr, 1), *r = palloc(pr, 1), *tmp = palloc(pr, 1);
    uint64_t *tdig = palloc(pr, ell);
    memcpy(a_ntt, a_ct, (size_t)gamma * P * 8);                       /* :941-949 */
    for (uint32_t j = 0; j < gamma; j++) or_ntt_forward(c, a_ntt + (size_t)j * P);
    for (uint32_t i = 0; i < gamma; i++) {                            /* step 1, :952-982 */
        memset(r, 0, P * 8);
        for (uint32_t j = 0; j < gamma; j++)  /* index = j * gen_pows[(n - i) % n] mod 2n */
            poly_mul_acc(pr, c, r, pk->mono + ((size_t)i * gamma + j) * P, a_ntt + (size_t)j * P);
        poly_pointwise(pr, c, r, pk->mod_inv_ntt, tmp);
        pack_automorph_ntt(pk, tmp, pk->tables + (size_t)i * n, r_all + (size_t)i * P);
    }
    for (int i = (int)gamma - 2; i >= 0; i--) {                       /* step 2, :987-1008 */
        memcpy(tmp, r_all + (size_t)(i + 1) * P, P * 8);
        or_ntt_inverse(c, tmp);
        or_gadget_decompose(pr, tmp, tdig);
        for (uint32_t k = 0; k < ell; k++) {
            memcpy(mono, tdig + (size_t)k * P, P * 8);
            or_ntt_forward(c, mono);
            poly_mul_acc(pr, c, r_all + (size_t)i * P, mono, pk->w_all_ntt + ((size_t)i * ell + k) * P);
        }
        memcpy(bold_t + (size_t)i * ell * P, tdig, (size_t)ell * P * 8);
    }
    memcpy(a_hat, r_all, P * 8);                                      /* :1029-1030 */
    or_ntt_inverse(c, a_hat);
    free(a_ntt); free(r_all); free(mono); free(r); free(tmp); free(tdig);
}

void or_generate_rotations(const or_pack *pk, const uint64_t *y_body, uint64_t *y_all) { /* :1309-1323 */
    const size_t P = psz(&pk->pr);
    const uint32_t ell = pk->pr.gadget_len;
    for (uint32_t i = 0; i + 1 < pk->gamma; i++)
        for (uint32_t k = 0; k < ell; k++)
            or_apply_automorphism(&pk->pr, y_body + (size_t)k * P, pk->gen_pows[i],
                                  y_all + ((size_t)i * ell + k) * P);
}

void or_packing_online(const or_pack *pk, const uint64_t *a_hat, const uint64_t *bold_t,
                       const uint64_t *y_all, uint32_t y_all_len, const uint64_t *b_poly,
                       uint64_t *ct_out) { /* inspiring2.rs:1185-1222 */
    const or_params *pr = &pk->pr;
    const or_ntt *c = pk->c;
    const uint32_t ell = pr->gadget_len;
    const size_t P = psz(pr);
    uint64_t *sum = palloc(pr, 1), *yn = palloc(pr, 1), *tn = palloc(pr, 1);
    for (uint32_t i = 0; i + 1 < pk->gamma; i++) {
        if (i >= y_all_len) break;
        for (uint32_t k = 0; k < ell; k++) {
            memcpy(yn, y_all + ((size_t)i * ell + k) * P, P * 8);
            or_ntt_forward(c, yn);
            memcpy(tn, bold_t + ((size_t)i * ell + k) * P, P * 8);
            or_ntt_forward(c, tn);
            poly_mul_acc(pr, c, sum, yn, tn); /* mul_ntt_domain + add_ntt_domain */
        }
    }
    or_ntt_inverse(c, sum);
    memcpy(ct_out, a_hat, P * 8);
    poly_add(pr, b_poly, sum, ct_out + P);
    free(sum); free(yn); free(tn);
}

int or_respond_inspiring(const or_params *pr, const or_ntt *c, const or_pack *pk,
                         const uint64_t *shard_polys, uint32_t ncols, const uint64_t *rgsw,
                         const uint64_t *y_body, const uint64_t *y_all, uint32_t y_all_len,
                         uint64_t *ct_out) { /* respond.rs:284-386 */
    const size_t P = psz(pr);
    const uint32_t d = pr->d, ell = pr->gadget_len;
    if (ncols == 0) { memset(ct_out, 0, 2 * P * 8); return 1; }         /* :307-313 */
    uint64_t *cols = palloc(pr, 2 * (size_t)ncols);
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t col = 0; col < ncols; col++) {                       /* :316-324 */
        uint64_t *rl = palloc(pr, 2);
        trivial_encrypt(pr, shard_polys + (size_t)col * P, rl);
        or_external_product(pr, c, rl, rgsw, cols + (size_t)col * 2 * P, 0);
        free(rl);
    }
    uint64_t *a_ct = palloc(pr, ncols), *bcf = (uint64_t *)calloc(d, sizeof(uint64_t));
    uint64_t *b_poly = palloc(pr, 1);
    const uint64_t inv = crt_inv(pr);
    for (uint32_t col = 0; col < ncols; col++) {
        memcpy(a_ct + (size_t)col * P, cols + (size_t)col * 2 * P, P * 8); /* :344-347 */
        if (col < d) bcf[col] = poly_coeff_inv(pr, cols + (size_t)col * 2 * P + P, 0, inv); /* :350-356 */
    }
    or_poly_from_coeffs(pr, bcf, b_poly);
    int rc = 0;
    if (y_all_len == 0 && y_body == NULL) {                           /* :366-369 */
        rc = -3;
    } else {
        uint64_t *a_hat = palloc(pr, 1);
        uint32_t nrot = ncols - 1;
        uint64_t *bold_t = palloc(pr, (size_t)(nrot ? nrot : 1) * ell);
        or_packing_offline(pk, a_ct, a_hat, bold_t);                  /* :360-362 */
        uint64_t *ya = NULL;
        const uint64_t *yuse = y_all;
        uint32_t ylen = y_all_len;
        if (y_all_len == 0) {                                          /* :365-377 */
            ya = palloc(pr, (size_t)(nrot ? nrot : 1) * ell);
            or_generate_rotations(pk, y_body, ya);
            yuse = ya;
            ylen = nrot;
        }
        or_packing_online(pk, a_hat, bold_t, yuse, ylen, b_poly, ct_out); /* :380 */
        free(a_hat); free(bold_t); free(ya);
    }
    free(cols); free(a_ct); free(bcf); free(b_poly);
    return rc;
}

/* n independent respond_inspiring calls, one query per thread (the CPU
 * baseline's throughput form: rayon over queries instead of over columns). */
void or_respond_inspiring_many(const or_params *pr, const or_ntt *c, const or_pack *pk, uint32_t n,
                               const uint64_t *const *shards, uint32_t ncols, const uint64_t *const *rgsw,
                               const uint64_t *const *y_body, uint64_t *const *ct_out, int *rc) {
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t i = 0; i < n; i++)
        rc[i] = or_respond_inspiring(pr, c, pk, shards[i], ncols, rgsw[i], y_body[i], NULL, 0, ct_out[i]);
}

/* respond_inspiring as the reference runs it (respond.rs:284-386), nothing
 * hoisted, for the CPU baseline's literal form:
 *  - every column's external product under an NTT context of its own
 *    (crs.params.ntt_context(), respond.rs:319-321; ntt.rs:116-164) with one
 *    mul_ntt per row product (external_product.rs:122-150);
 *  - PackParams::new in full (inspiring2.rs:107-182) under its own context:
 *    the n generator powers, mod_inv's NTT, the 4n monomials (X^j and -X^j,
 *    coefficient and NTT form) and the brute-force NTT-domain automorphism
 *    table of every odd t < 2n (inspiring2.rs:208-262); packing reads its
 *    tables and monomials out of these;
 *  - OfflinePackingKeys::generate (inspiring2.rs:448-506), packing_offline,
 *    generate_rotations and packing_online under the query's context
 *    (respond.rs:293).
 * One thread.  Returns or_respond_inspiring's code (and the same ciphertext),
 * or -4 when a brute-force table disagrees with the closed form.  work[0..1]
 * (may be NULL) receive the brute-force retries and the tables built. */
int or_respond_inspiring_literal(const or_params *pr, const uint8_t w_seed[32], const uint64_t *shard_polys,
                                 uint32_t ncols, const uint64_t *rgsw, const uint64_t *y_body, uint64_t *ct_out,
                                 uint64_t *work) {
    const size_t P = psz(pr);
    const uint32_t d = pr->d, n = pr->d, ell = pr->gadget_len;
    if (ncols == 0) { memset(ct_out, 0, 2 * P * 8); return 1; }
    or_ntt *ctx = or_ntt_new(d, pr->moduli, pr->K);                      /* :293 */
    uint64_t *cols = palloc(pr, 2 * (size_t)ncols);
    for (uint32_t col = 0; col < ncols; col++) {                         /* :316-324 */
        or_ntt *local = or_ntt_new(d, pr->moduli, pr->K);
        uint64_t *rl = palloc(pr, 2);
        trivial_encrypt(pr, shard_polys + (size_t)col * P, rl);
        or_external_product(pr, local, rl, rgsw, cols + (size_t)col * 2 * P, 1);
        free(rl);
        or_ntt_free(local);
    }
    uint64_t *a_ct = palloc(pr, ncols), *bcf = (uint64_t *)calloc(d, sizeof(uint64_t));
    uint64_t *b_poly = palloc(pr, 1);
    const uint64_t inv = crt_inv(pr);
    for (uint32_t col = 0; col < ncols; col++) {
        memcpy(a_ct + (size_t)col * P, cols + (size_t)col * 2 * P, P * 8); /* :344-347 */
        if (col < d) bcf[col] = poly_coeff_inv(pr, cols + (size_t)col * 2 * P + P, 0, inv); /* :350-356 */
    }
    or_poly_from_coeffs(pr, bcf, b_poly);

    /* PackParams::new(&crs.params, num_columns) (:360) */
    or_ntt *pctx = or_ntt_new(d, pr->moduli, pr->K);                     /* inspiring2.rs:111 */
    or_pack *pk = pack_alloc(pr, pctx, ncols, 0);
    uint64_t *mono = palloc(pr, n), *mono_ntt = palloc(pr, n);           /* :139-158 */
    uint64_t *neg = palloc(pr, n), *neg_ntt = palloc(pr, n);
    uint64_t *cf = (uint64_t *)calloc(n, sizeof(uint64_t));
    for (uint32_t j = 0; j < n; j++) {
        memset(cf, 0, (size_t)n * 8);
        cf[j] = 1;
        or_poly_from_coeffs(pr, cf, mono + (size_t)j * P);
        memcpy(mono_ntt + (size_t)j * P, mono + (size_t)j * P, P * 8);
        or_ntt_forward(pctx, mono_ntt + (size_t)j * P);
        cf[j] = pr->q - 1;
        or_poly_from_coeffs(pr, cf, neg + (size_t)j * P);
        memcpy(neg_ntt + (size_t)j * P, neg + (size_t)j * P, P * 8);
        or_ntt_forward(pctx, neg_ntt + (size_t)j * P);
    }
    free(cf);
    uint32_t *all = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n * n); /* :208-262, table (t-1)/2 */
    uint64_t retries = 0;
    for (uint32_t t = 1; t < 2 * n; t += 2)
        retries += (uint64_t)or_automorph_table_bruteforce(pr, pctx, t, 0x5eed0000ull + t, all + (size_t)table_index(t) * n);
    int bad = 0;
    uint32_t *closed = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < ncols; i++) {
        const uint32_t *tab = all + (size_t)table_index(pk->gen_pows[i]) * n;
        or_automorph_table(n, 2 * table_index(pk->gen_pows[i]) + 1, closed);
        if (memcmp(tab, closed, sizeof(uint32_t) * n)) bad = 1;
        memcpy(pk->tables + (size_t)i * n, tab, sizeof(uint32_t) * n);
        for (uint32_t j = 0; j < ncols; j++) {                           /* packing_offline's index, :955-967 */
            const uint32_t e = (uint32_t)(((uint64_t)j * pk->gen_pows[(n - i) % n]) % (2ull * n));
            memcpy(pk->mono + ((size_t)i * ncols + j) * P, (e < n ? mono_ntt : neg_ntt) + (size_t)(e % n) * P, P * 8);
        }
    }
    free(closed);
    pk->c = ctx;  /* packing runs under the query's context (:293, :361) */
    pack_keys(pk, w_seed);                                               /* :361 */
    if (work) { work[0] = retries; work[1] = n; }
    int rc = 0;
    if (bad) {
        rc = -4;
    } else if (y_body == NULL) {
        rc = -3;
    } else {
        uint64_t *a_hat = palloc(pr, 1);
        uint32_t nrot = ncols - 1;
        uint64_t *bold_t = palloc(pr, (size_t)(nrot ? nrot : 1) * ell);
        uint64_t *ya = palloc(pr, (size_t)(nrot ? nrot : 1) * ell);
        or_packing_offline(pk, a_ct, a_hat, bold_t);                     /* :362 */
        or_generate_rotations(pk, y_body, ya);                           /* :365-377 */
        or_packing_online(pk, a_hat, bold_t, ya, nrot, b_poly, ct_out);  /* :380 */
        free(a_hat); free(bold_t); free(ya);
    }
    or_pack_free(pk);
    free(mono); free(mono_ntt); free(neg); free(neg_ntt); free(all);
    or_ntt_free(pctx);
    or_ntt_free(ctx);
    free(cols); free(a_ct); free(bcf); free(b_poly);
    return rc;
}

/* n literal respond_inspiring calls, one query per thread. */
void or_respond_inspiring_literal_many(const or_params *pr, const uint8_t w_seed[32], uint32_t n,
                                       const uint64_t *const *shards, uint32_t ncols, const uint64_t *const *rgsw,
                                       const uint64_t *const *y_body, uint64_t *const *ct_out, int *rc) {
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t i = 0; i < n; i++)
        rc[i] = or_respond_inspiring_literal(pr, w_seed, shards[i], ncols, rgsw[i], y_body[i], ct_out[i], NULL);
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Tree packing: inspiring/automorph_pack.rs, rlwe/galois.rs            */
/* ------------------------------------------------------------------ */

/* RlweCiphertext::add (rlwe/enc.rs:114-119) */
static void ct_add(const or_params *pr, const uint64_t *x, const uint64_t *y, uint64_t *out) {
    poly_add(pr, x, y, out);
    poly_add(pr, x + psz(pr), y + psz(pr), out + psz(pr));
}

static int ct_is_zero(const or_params *pr, const uint64_t *ct) {
    for (size_t x = 0; x < 2 * psz(pr); x++)
        if (ct[x]) return 0;
    return 1;
}

/* sample_extract_coeff0 (rlwe/enc.rs:195-208): a'_0 = a_0, a'_i = a_(d-i),
 * b = b_0, all as composite values. */
void or_sample_extract_coeff0(const or_params *pr, const uint64_t *ct, uint64_t *a_out, uint64_t *b_out) {
    const uint32_t d = pr->d;
    const uint64_t inv = crt_inv(pr);
    a_out[0] = poly_coeff_inv(pr, ct, 0, inv);
    for (uint32_t i = 1; i < d; i++) a_out[i] = poly_coeff_inv(pr, ct, d - i, inv);
    *b_out = poly_coeff_inv(pr, ct + psz(pr), 0, inv);
}

/* homomorphic_automorph (automorph_pack.rs:44-78): tau_t on both parts
 * (galois.rs:64-69), gadget digits of tau_t(a), then
 * a' = sum_i D_i * ks[i].a and b' = tau_t(b) + sum_i D_i * ks[i].b, each
 * product through mul_ntt (poly.rs:505-532). */
void or_homomorphic_automorph(const or_params *pr, const or_ntt *c, const uint64_t *ct, uint32_t t,
                              const uint64_t *ks, uint32_t ks_rows, uint64_t *out) {
    const size_t P = psz(pr);
    const uint32_t ell = pr->gadget_len;
    uint64_t *ta = palloc(pr, 2), *dig = palloc(pr, ell), *tm = palloc(pr, 1);
    or_apply_automorphism(pr, ct, t, ta);
    or_apply_automorphism(pr, ct + P, t, ta + P);
    or_gadget_decompose(pr, ta, dig);
    uint64_t *ra = out, *rb = out + P;
    memset(ra, 0, P * 8);
    memcpy(rb, ta + P, P * 8);
    for (uint32_t i = 0; i < ell && i < ks_rows; i++) {
        const uint64_t *row = ks + (size_t)i * 2 * P;
        or_poly_mul_ntt(pr, c, dig + (size_t)i * P, row, tm);
        poly_add(pr, ra, tm, ra);
        or_poly_mul_ntt(pr, c, dig + (size_t)i * P, row + P, tm);
        poly_add(pr, rb, tm, rb);
    }
    free(ta); free(dig); free(tm);
}

typedef struct {
    const or_params *pr;
    const or_ntt *c;
    const uint64_t *leaves;    /* [d][2][K][d] prepped RLWEs */
    const uint8_t *leaf_nz;    /* leaf is nonzero; NULL in the literal recursion */
    const uint64_t *keys;      /* galois keys [nkeys][ell][2][K][d] */
    uint32_t nkeys, log_n;
    const uint64_t *y, *neg_y; /* [log_d][K][d] YConstants (automorph_pack.rs:86-148) */
} tree_ctx;

/* Returns 1 when the node's ciphertext is exactly zero and was left unset
 * (only when leaf_nz is given: a subtree of zero leaves packs to zero, since
 * every step is linear with exact zero inputs). */
static int pack_inner(const tree_ctx *T, uint32_t ell, uint32_t start, uint64_t *out) { /* :171-256 */
    const or_params *pr = T->pr;
    const size_t P = psz(pr), CT = 2 * P;
    if (ell == 0) {
        if (T->leaf_nz && !T->leaf_nz[start]) return 1;
        memcpy(out, T->leaves + (size_t)start * CT, CT * 8);
        return 0;
    }
    uint32_t step = 1u << (T->log_n - ell);
    uint64_t *ev = palloc(pr, 2), *od = palloc(pr, 2);
    int ez = pack_inner(T, ell - 1, start, ev);
    int oz = pack_inner(T, ell - 1, start + step, od);
    if (ez && oz) { free(ev); free(od); return 1; }
    if (ez) memset(ev, 0, CT * 8);
    if (oz) memset(od, 0, CT * 8);
    const uint64_t *y = T->y + (size_t)(ell - 1) * P, *ny = T->neg_y + (size_t)(ell - 1) * P;
    uint64_t *s0 = palloc(pr, 2), *s1 = palloc(pr, 2), *tm = palloc(pr, 2);
    /* ct_sum_0 = even + y*odd, ct_sum_1 = even + (-y)*odd (RlweCiphertext::poly_mul, rlwe/enc.rs:147-152) */
    or_poly_mul_ntt(pr, T->c, od, y, tm);
    or_poly_mul_ntt(pr, T->c, od + P, y, tm + P);
    ct_add(pr, ev, tm, s0);
    or_poly_mul_ntt(pr, T->c, od, ny, tm);
    or_poly_mul_ntt(pr, T->c, od + P, ny, tm + P);
    ct_add(pr, ev, tm, s1);
    uint32_t t = (1u << ell) + 1, ks_idx = T->nkeys - ell;
    const size_t KS = (size_t)pr->gadget_len * CT;
    or_homomorphic_automorph(pr, T->c, s1, t, T->keys + (size_t)ks_idx * KS, pr->gadget_len, tm);
    ct_add(pr, s0, tm, out);
    free(ev); free(od); free(s0); free(s1); free(tm);
    return 0;
}

/* pack_lwes (automorph_pack.rs:418-487) of n LWEs: lwe_a [n][d] and lwe_b[n]
 * composite values as sample_extract_coeff0 gives them.  ct_out receives a
 * [K][d] followed by b [K][bdim]; the return value is bdim.  With two moduli
 * the reference adds b_val*d to the first d of the 2d raw residue words of b
 * and rebuilds b from all 2d words as composite coefficients
 * (Poly::from_coeffs_moduli, poly.rs:127-159), so bdim = 2d; residue_wise = 1
 * adds b_val*d to each residue instead (bdim = d, a corrected mode). */
uint32_t or_pack_lwes(const or_params *pr, const or_ntt *c, uint32_t n, const uint64_t *lwe_a,
                      const uint64_t *lwe_b, const uint64_t *galois_keys, uint32_t nkeys, int literal,
                      int residue_wise, uint64_t *ct_out) {
    const uint32_t d = pr->d, K = pr->K;
    const size_t P = psz(pr), CT = 2 * P;
    const uint64_t q = pr->q;
    if (n == 0) { memset(ct_out, 0, CT * 8); return d; }               /* :426-431 */
    uint64_t *cf = (uint64_t *)calloc(2 * (size_t)d, sizeof(uint64_t));
    uint32_t m = n < d ? n : d;
    uint64_t *leaves = palloc(pr, 2 * (size_t)d);                        /* padded to d (:448-454) */
    uint8_t *nz = (uint8_t *)calloc(d, 1);
    for (uint32_t j = 0; j < m; j++) {                                   /* prep_pack_lwes (:383-405) */
        const uint64_t *a = lwe_a + (size_t)j * d;
        cf[0] = a[0];                                                    /* invert_sample_extract (:349-363) */
        for (uint32_t i = 1; i < d; i++) cf[i] = a[d - i];
        or_poly_from_coeffs(pr, cf, leaves + (size_t)j * CT);
        nz[j] = !ct_is_zero(pr, leaves + (size_t)j * CT);
    }
    if (n == 1) {                                                        /* :436-443 */
        memcpy(ct_out, leaves, P * 8);
        memset(cf, 0, (size_t)d * 8);
        cf[0] = lwe_b[0];
        or_poly_from_coeffs(pr, cf, ct_out + P);
        free(cf); free(leaves); free(nz);
        return d;
    }
    uint32_t log_d = 0;
    while ((1u << log_d) < d) log_d++;
    uint64_t *y = palloc(pr, log_d), *ny = palloc(pr, log_d);
    for (uint32_t l = 0; l < log_d; l++) {                               /* YConstants::generate (:104-131) */
        uint32_t st = d >> (l + 1);
        memset(cf, 0, (size_t)d * 8);
        cf[st] = 1;
        or_poly_from_coeffs(pr, cf, y + (size_t)l * P);
        cf[st] = q - 1;
        or_poly_from_coeffs(pr, cf, ny + (size_t)l * P);
    }
    tree_ctx T = {pr, c, leaves, literal ? NULL : nz, galois_keys, nkeys, log_d, y, ny};
    uint64_t *packed = palloc(pr, 2);
    if (pack_inner(&T, log_d, 0, packed)) memset(packed, 0, CT * 8);
    memcpy(ct_out, packed, P * 8);
    uint32_t bdim = d;
    if (residue_wise || K == 1) {
        for (uint32_t k = 0; k < K; k++) {
            uint64_t qk = pr->moduli[k];
            for (uint32_t z = 0; z < d; z++) {
                uint64_t sc = z < m ? (uint64_t)(((u128)lwe_b[z] * d) % q) : 0;
                ct_out[P + (size_t)k * d + z] = (uint64_t)(((u128)packed[P + (size_t)k * d + z] + sc % qk) % qk);
            }
        }
    } else {                                                             /* :476-484 */
        bdim = 2 * d;
        uint64_t *raw = packed + P;                                      /* b.coeffs(): [res0 | res1] */
        for (uint32_t z = 0; z < d; z++) {
            uint64_t sc = z < m ? (uint64_t)(((u128)lwe_b[z] * d) % q) : 0; /* ModQ::mul */
            raw[z] = (uint64_t)(((u128)raw[z] + sc) % q);                    /* ModQ::add */
        }
        for (uint32_t k = 0; k < K; k++)
            for (uint32_t i = 0; i < 2 * d; i++) ct_out[P + (size_t)k * 2 * d + i] = raw[i] % pr->moduli[k];
    }
    free(cf); free(leaves); free(nz); free(y); free(ny); free(packed);
    return bdim;
}

/* respond_one_packing (respond.rs:213-268): columns by external product, an
 * LWE from coefficient 0 of each, pack_lwes with the CRS galois keys.
 * Returns bdim (see or_pack_lwes), or 0 for the empty shard (ct_out = the
 * zero ciphertext; the reference also returns column_ciphertexts = [0]). */
uint32_t or_respond_one_packing(const or_params *pr, const or_ntt *c, const uint64_t *shard_polys,
                                uint32_t ncols, const uint64_t *rgsw, const uint64_t *galois_keys,
                                uint32_t nkeys, int literal, int residue_wise, uint64_t *ct_out) {
    const size_t P = psz(pr);
    const uint32_t d = pr->d;
    if (ncols == 0) { memset(ct_out, 0, 2 * P * 8); return 0; }          /* :230-236 */
    uint64_t *cols = palloc(pr, 2 * (size_t)ncols);
#pragma omp parallel for schedule(dynamic, 1) if (!literal)
    for (uint32_t col = 0; col < ncols; col++) {                         /* :241-249 */
        uint64_t *rl = palloc(pr, 2);
        trivial_encrypt(pr, shard_polys + (size_t)col * P, rl);
        or_external_product(pr, c, rl, rgsw, cols + (size_t)col * 2 * P, literal);
        free(rl);
    }
    uint64_t *la = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)ncols * d);
    uint64_t *lb = (uint64_t *)malloc(sizeof(uint64_t) * ncols);
    for (uint32_t col = 0; col < ncols; col++)                            /* :253-256 */
        or_sample_extract_coeff0(pr, cols + (size_t)col * 2 * P, la + (size_t)col * d, lb + col);
    uint32_t bdim = or_pack_lwes(pr, c, ncols, la, lb, galois_keys, nkeys, literal, residue_wise, ct_out);
    free(cols); free(la); free(lb);
    return bdim;
}

/* n respond_one_packing calls, one query per thread. */
void or_respond_one_packing_many(const or_params *pr, const or_ntt *c, uint32_t n,
                                 const uint64_t *const *shards, uint32_t ncols, const uint64_t *const *rgsw,
                                 const uint64_t *galois_keys, uint32_t nkeys, uint64_t *const *ct_out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t i = 0; i < n; i++)
        or_respond_one_packing(pr, c, shards[i], ncols, rgsw[i], galois_keys, nkeys, 0, 0, ct_out[i]);
}

/* ------------------------------------------------------------------ */
/* Client side: fixtures and decrypt checks only                        */
/* ------------------------------------------------------------------ */

/* A rounded normal sample from the ChaCha stream (Box-Muller).  The
 * reference's GaussianSampler (math/gaussian.rs) is client-side and its draws
 * are not part of the respond contract: queries are inputs to respond. */
static int64_t gauss_sample(or_chacha *r, double sigma) {
    double u1 = ((double)(or_chacha_next_u64(r) >> 11) + 1.0) / 9007199254740993.0;
    double u2 = (double)(or_chacha_next_u64(r) >> 11) / 9007199254740992.0;
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2) * sigma;
    return (int64_t)llround(z);
}

static uint64_t from_signed(int64_t v, uint64_t q) { /* math/modular.rs:110-117 */
    if (v >= 0) return (uint64_t)v % q;
    uint64_t a = (uint64_t)(-v);
    return q - (a % q);
}

void or_sample_gaussian_poly(const or_params *pr, or_chacha *r, double sigma, uint64_t *out) {
    /* Poly::sample_gaussian_moduli (poly.rs:208-232): one sample per coefficient, all residues */
    for (uint32_t i = 0; i < pr->d; i++) {
        int64_t s = gauss_sample(r, sigma);
        for (uint32_t k = 0; k < pr->K; k++) out[(size_t)k * pr->d + i] = from_signed(s, pr->moduli[k]);
    }
}

void or_rgsw_encrypt(const or_params *pr, const or_ntt *c, const uint64_t *sk, const uint64_t *msg,
                     or_chacha *r, double sigma, uint64_t *rgsw) { /* rgsw/types.rs:133-187 */
    const size_t P = psz(pr);
    const uint32_t ell = pr->gadget_len;
    uint64_t pw[16];
    gadget_powers(pr, pw);
    uint64_t *arand = palloc(pr, 1), *e = palloc(pr, 1), *as = palloc(pr, 1), *sm = palloc(pr, 1);
    for (uint32_t half = 0; half < 2; half++)
        for (uint32_t i = 0; i < ell; i++) {
            uint64_t *ra = rgsw + (size_t)(2 * (half * ell + i)) * P, *rb = ra + P;
            poly_random_rng(pr, r, arand);
            or_sample_gaussian_poly(pr, r, sigma, e);
            or_poly_mul_ntt(pr, c, arand, sk, as);
            poly_neg(pr, as, as);
            poly_add(pr, as, e, rb);                    /* b = -a s + e */
            poly_scalar_mul(pr, msg, pw[i], sm);
            if (half == 0) { poly_add(pr, arand, sm, ra); }
            else { memcpy(ra, arand, P * 8); poly_add(pr, rb, sm, rb); }
        }
    free(arand); free(e); free(as); free(sm);
}

void or_seeded_rgsw_encrypt(const or_params *pr, const or_ntt *c, const uint64_t *sk,
                            const uint64_t *msg, or_chacha *r, double sigma, uint8_t *seeds,
                            uint64_t *b_out) { /* rgsw/types.rs:235-300 */
    const size_t P = psz(pr);
    const uint32_t ell = pr->gadget_len;
    uint64_t pw[16];
    gadget_powers(pr, pw);
    uint64_t *a = palloc(pr, 1), *e = palloc(pr, 1), *as = palloc(pr, 1), *sm = palloc(pr, 1);
    for (uint32_t half = 0; half < 2; half++)
        for (uint32_t i = 0; i < ell; i++) {
            uint32_t row = half * ell + i;
            uint8_t *sd = seeds + 32 * row;
            for (int w = 0; w < 8; w++) {
                uint32_t x = or_chacha_next_u32(r);
                memcpy(sd + 4 * w, &x, 4);
            }
            uint64_t *b = b_out + (size_t)row * P;
            or_poly_from_seed(pr, sd, a);
            or_sample_gaussian_poly(pr, r, sigma, e);
            or_poly_mul_ntt(pr, c, a, sk, as);
            poly_neg(pr, as, as);
            poly_add(pr, as, e, b);
            poly_scalar_mul(pr, msg, pw[i], sm);
            if (half == 0) { or_poly_mul_ntt(pr, c, sm, sk, as); poly_add(pr, b, as, b); }
            else poly_add(pr, b, sm, b);
        }
    free(a); free(e); free(as); free(sm);
}

void or_inverse_monomial(const or_params *pr, uint32_t k, uint64_t *out) { /* encode_db.rs:80-91 */
    uint64_t *cf = (uint64_t *)calloc(pr->d, sizeof(uint64_t));
    if (k == 0) cf[0] = 1;
    else cf[pr->d - k] = pr->q - 1;
    or_poly_from_coeffs(pr, cf, out);
    free(cf);
}

/* generate_ksk_body (inspiring2.rs:846-904) */
static void generate_ksk_body(const or_params *pr, const or_ntt *c, const uint64_t *s, uint32_t gen,
                              const uint64_t *w_mask, or_chacha *r, double sigma, uint64_t *out) {
    const size_t P = psz(pr);
    uint64_t pw[16];
    gadget_powers(pr, pw);
    uint64_t *tau_s = palloc(pr, 1), *t1 = palloc(pr, 1), *t2 = palloc(pr, 1), *e = palloc(pr, 1);
    uint64_t *ecf = (uint64_t *)malloc(sizeof(uint64_t) * pr->d);
    or_apply_automorphism(pr, s, gen, tau_s);
    for (uint32_t k = 0; k < pr->gadget_len; k++) {
        poly_scalar_mul(pr, tau_s, pw[k], t1);
        or_poly_mul_ntt(pr, c, s, w_mask + (size_t)k * P, t2);
        poly_neg(pr, t2, t2);
        for (uint32_t i = 0; i < pr->d; i++) {
            int64_t smp = gauss_sample(r, sigma);
            ecf[i] = smp >= 0 ? (uint64_t)smp % pr->q
                              : (uint64_t)((int64_t)pr->q + (smp % (int64_t)pr->q));
        }
        or_poly_from_coeffs(pr, ecf, e);
        poly_add(pr, t1, t2, out + (size_t)k * P);
        poly_add(pr, out + (size_t)k * P, e, out + (size_t)k * P);
    }
    free(tau_s); free(t1); free(t2); free(e); free(ecf);
}

/* ClientPackingKeys::generate (inspiring2.rs:640-707): y_body [ell][K][d] and
 * the in-process y_all [(gamma-1)][ell][K][d] built through NTT-domain rotations. */
void or_client_packing_keys(const or_params *pr, const or_ntt *c, const or_pack *pk,
                            const uint64_t *sk, or_chacha *r, double sigma, uint64_t *y_body,
                            uint64_t *y_all) {
    const size_t P = psz(pr);
    const uint32_t ell = pr->gadget_len, n = pr->d;
    generate_ksk_body(pr, c, sk, pk->gen_pows[1], pk->w_mask, r, sigma, y_body);
    if (y_all) {
        uint64_t *yn = palloc(pr, 1);
        for (uint32_t i = 0; i + 1 < pk->gamma; i++)
            for (uint32_t k = 0; k < ell; k++) {
                memcpy(yn, y_body + (size_t)k * P, P * 8);
                or_ntt_forward(c, yn);
                uint64_t *dst = y_all + ((size_t)i * ell + k) * P;
                pack_automorph_ntt(pk, yn, pk->tables + (size_t)i * n, dst);
                or_ntt_inverse(c, dst);
            }
        free(yn);
    }
}

void or_decrypt(const or_params *pr, const or_ntt *c, const uint64_t *ct, const uint64_t *sk,
                uint64_t *msg) { /* rlwe/enc.rs:90-108 */
    const size_t P = psz(pr);
    uint64_t *as = palloc(pr, 1);
    or_poly_mul_ntt(pr, c, ct, sk, as);
    poly_add(pr, as, ct + P, as);
    const uint64_t delta = or_delta(pr), half = delta / 2, inv = crt_inv(pr);
    for (uint32_t i = 0; i < pr->d; i++) {
        uint64_t v = poly_coeff_inv(pr, as, i, inv);
        uint64_t rounded = (uint64_t)(((u128)v + half) / delta);
        msg[i] = rounded % pr->p;
    }
    free(as);
}

static uint64_t extract_column_value(const uint8_t *e, uint32_t entry_size, uint32_t col) { /* encode_db.rs:169-181 */
    uint32_t off = col * 2;
    if (off + 1 < entry_size) return (uint64_t)e[off] | ((uint64_t)e[off + 1] << 8);
    if (off < entry_size) return e[off];
    return 0;
}

void or_encode_shard(const or_params *pr, const uint8_t *db, uint64_t num_entries, uint32_t entry_size,
                     uint64_t entries_per_shard, uint64_t shard_idx, uint64_t *polys) { /* encode_db.rs:106-164 */
    uint32_t npolys = (entry_size * 8 + 15) / 16;
    uint64_t off = shard_idx * entries_per_shard;
    uint64_t actual = num_entries - off < entries_per_shard ? num_entries - off : entries_per_shard;
    uint64_t *cf = (uint64_t *)malloc(sizeof(uint64_t) * pr->d);
    for (uint32_t pi = 0; pi < npolys; pi++) {
        memset(cf, 0, sizeof(uint64_t) * pr->d);
        for (uint64_t li = 0; li < actual && li < pr->d; li++)
            cf[li] = extract_column_value(db + (off + li) * entry_size, entry_size, pi) % pr->q;
        or_poly_from_coeffs(pr, cf, polys + (size_t)pi * psz(pr));
    }
    free(cf);
}

/* generate_automorphism_ks_matrix (ks/setup.rs:218-247) through
 * generate_ks_matrix (ks/setup.rs:118-166): from tau_t(s) to s, rows
 * (a_i, -a_i s + e_i + tau_t(s) z^i), out [ell][2][K][d].  The draws come from
 * the ChaCha stream r (the reference samples a with thread_rng). */
void or_automorphism_ks_matrix(const or_params *pr, const or_ntt *c, const uint64_t *sk, uint32_t t,
                               or_chacha *r, double sigma, uint64_t *out) {
    const size_t P = psz(pr);
    uint64_t pw[16];
    gadget_powers(pr, pw);
    uint64_t *ts = palloc(pr, 1), *e = palloc(pr, 1), *as = palloc(pr, 1), *sm = palloc(pr, 1);
    or_apply_automorphism(pr, sk, t, ts);
    for (uint32_t i = 0; i < pr->gadget_len; i++) {
        uint64_t *ra = out + (size_t)i * 2 * P, *rb = ra + P;
        poly_random_rng(pr, r, ra);
        or_sample_gaussian_poly(pr, r, sigma, e);
        or_poly_mul_ntt(pr, c, ra, sk, as);
        poly_neg(pr, as, as);
        poly_add(pr, as, e, rb);
        poly_scalar_mul(pr, ts, pw[i], sm);
        poly_add(pr, rb, sm, rb);
    }
    free(ts); free(e); free(as); free(sm);
}

/* ServerCrs.galois_keys (setup.rs:167-173): for i < log_d, the matrix of
 * t = d/2^i + 1; out [log_d][ell][2][K][d]. */
void or_galois_keys(const or_params *pr, const or_ntt *c, const uint64_t *sk, or_chacha *r, double sigma,
                    uint64_t *out) {
    uint32_t log_d = 0;
    while ((1u << log_d) < pr->d) log_d++;
    const size_t KS = (size_t)pr->gadget_len * 2 * psz(pr);
    for (uint32_t i = 0; i < log_d; i++)
        or_automorphism_ks_matrix(pr, c, sk, (pr->d >> i) + 1, r, sigma, out + (size_t)i * KS);
}
