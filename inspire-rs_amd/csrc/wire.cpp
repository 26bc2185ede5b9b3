// Wire formats of the respond path: ServerResponse out, ClientQuery /
// SeededClientQuery in, as the reference's server and clients exchange them.
//
//  bincode 1.3 (bincode::serialize, ServerResponse::to_binary,
//  src/pir/respond.rs:43-51): little-endian fixed-width integers, usize as
//  u64, Vec as a u64 length then its elements, [u8; 32] as 32 bytes, bool as
//  one byte, Option as a tag byte then the value, unit enum variants as a u32
//  index, struct fields in declaration order.  Fields marked
//  skip_serializing_if are left out entirely when skipped: ClientQuery's
//  inspiring_packing_keys when None (src/pir/query.rs:126-131) and
//  ClientPackingKeys.z_body when empty (src/inspiring/inspiring2.rs:609-610).
//  serde_json (the HTTP handlers, src/bin/server.rs:120-236): compact
//  objects with the same field names and order, integers in decimal,
//  PackingMode as "inspiring" / "tree" (rename_all = "snake_case").
//
//  Poly (src/math/poly.rs:63-77): coeffs (the residue-major words, K x dim),
//  moduli, q, dim, crt_q0_inv_mod_q1 (0 with one modulus, poly.rs:80-93),
//  is_ntt.  RlweCiphertext = (a, b); SeededRlweCiphertext = (seed, b);
//  RgswCiphertext / SeededRgswCiphertext = (rows, gadget); GadgetVector =
//  (base, len, q).
#include <charconv>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "abi.hpp"

using namespace inspire;

// The parameter set as the formats need it (no device state).
struct WireParams {
    uint32_t K, N, ell;
    uint64_t moduli[2], q, inv, gadget_base;
};

struct inspire_wire_query {
    std::vector<uint64_t> rgsw;    // full: [2ell][2][K][d]; seeded: b polys [2ell][K][d]
    std::vector<uint8_t> seeds;    // seeded: [2ell][32]
    std::vector<uint64_t> y_body;  // [ell][K][d] when the query carries packing keys
};

namespace inspire {
namespace {

Error bad_wire(const std::string& m) { return Error{INSPIRE_ERR_INVALID_ARGUMENT, m}; }

// Poly::init_moduli (poly.rs:80-93): crt_q0_inv_mod_q1 = q0^-1 mod q1, or 0
// with one modulus.
uint64_t inv_mod(uint64_t a, uint64_t m) {
    int64_t t = 0, nt = 1;
    int64_t r = (int64_t)m, nr = (int64_t)(a % m);
    while (nr) {
        const int64_t qq = r / nr;
        int64_t x = t - qq * nt; t = nt; nt = x;
        x = r - qq * nr; r = nr; nr = x;
    }
    if (r != 1) throw Error{INSPIRE_ERR_INVALID_PARAMS, "CRT moduli must be coprime"};
    return (uint64_t)(t < 0 ? t + (int64_t)m : t);
}

WireParams wire_params(const inspire_params* p) {
    if (!p) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
    if (p->crt_count < 1 || p->crt_count > 2) throw Error{INSPIRE_ERR_INVALID_PARAMS, "crt_count must be 1 or 2"};
    if (!p->ring_dim || (p->ring_dim & (p->ring_dim - 1)) || p->ring_dim > (1u << 20))
        throw Error{INSPIRE_ERR_INVALID_PARAMS, "ring_dim must be a power of two"};
    // the readers index rows[0] and size every row array by 2 * gadget_len
    if (p->gadget_len == 0 || p->gadget_len > (uint32_t)kMaxEll)
        throw Error{INSPIRE_ERR_UNSUPPORTED, "gadget_len must be in [1, 4]"};
    WireParams w{};
    w.K = p->crt_count;
    w.N = p->ring_dim;
    w.ell = p->gadget_len;
    unsigned __int128 prod = 1;
    for (uint32_t k = 0; k < w.K; k++) {
        w.moduli[k] = p->crt_moduli[k];
        if (w.moduli[k] < 2) throw Error{INSPIRE_ERR_INVALID_PARAMS, "bad CRT modulus"};
        prod *= w.moduli[k];
    }
    if (prod != (unsigned __int128)p->q) throw Error{INSPIRE_ERR_INVALID_PARAMS, "q must equal product of CRT moduli"};
    w.q = p->q;
    w.inv = w.K == 2 ? inv_mod(w.moduli[0], w.moduli[1]) : 0;
    w.gadget_base = p->gadget_base;
    return w;
}

// ---------------------------------------------------------------------------
// Writers
// ---------------------------------------------------------------------------

struct Sink {
    virtual ~Sink() = default;
    virtual void bytes(const void* p, size_t n) = 0;
};

// Counts, then writes into a caller buffer of exactly that size.
struct BufSink : Sink {
    uint8_t* out = nullptr;
    size_t n = 0;
    void bytes(const void* p, size_t k) override {
        if (out) std::memcpy(out + n, p, k);
        n += k;
    }
};

struct Bin {
    Sink& s;
    void u8(uint8_t v) { s.bytes(&v, 1); }
    void u32(uint32_t v) { s.bytes(&v, 4); }  // little-endian host (x86-64)
    void u64(uint64_t v) { s.bytes(&v, 8); }
    void words(const uint64_t* p, size_t n) { s.bytes(p, n * 8); }
};

// Poly as bincode: dim coefficients per residue (dim = d, or 2d for a
// tree-packed b with two moduli).
void bin_poly(Bin& w, const WireParams* ctx, const uint64_t* residues, uint32_t dim) {
    const size_t n = (size_t)ctx->K * dim;
    w.u64(n);
    w.words(residues, n);
    w.u64(ctx->K);
    for (uint32_t k = 0; k < ctx->K; k++) w.u64(ctx->moduli[k]);
    w.u64(ctx->q);
    w.u64(dim);
    w.u64(ctx->inv);
    w.u8(0);  // is_ntt = false
}

struct Json {
    Sink& s;
    char buf[24];
    void raw(const char* t) { s.bytes(t, std::strlen(t)); }
    void num(uint64_t v) {
        auto r = std::to_chars(buf, buf + sizeof buf, v);
        s.bytes(buf, (size_t)(r.ptr - buf));
    }
};

void json_poly(Json& j, const WireParams* ctx, const uint64_t* residues, uint32_t dim) {
    const size_t n = (size_t)ctx->K * dim;
    j.raw("{\"coeffs\":[");
    // numbers in bulk: format into a local block, then one write
    char blk[4096];
    size_t used = 0;
    for (size_t i = 0; i < n; i++) {
        if (used > sizeof blk - 24) {
            j.s.bytes(blk, used);
            used = 0;
        }
        if (i) blk[used++] = ',';
        auto r = std::to_chars(blk + used, blk + sizeof blk, residues[i]);
        used = (size_t)(r.ptr - blk);
    }
    j.s.bytes(blk, used);
    j.raw("],\"moduli\":[");
    for (uint32_t k = 0; k < ctx->K; k++) {
        if (k) j.raw(",");
        j.num(ctx->moduli[k]);
    }
    j.raw("],\"q\":");
    j.num(ctx->q);
    j.raw(",\"dim\":");
    j.num(dim);
    j.raw(",\"crt_q0_inv_mod_q1\":");
    j.num(ctx->inv);
    j.raw(",\"is_ntt\":false}");
}

void check_response(const WireParams* ctx, const inspire_response* r) {
    if (!r || !r->ciphertext) throw bad_wire("response ciphertext buffer is null");
    const uint32_t bd = r->ciphertext_b_dim ? r->ciphertext_b_dim : ctx->N;
    if (bd != ctx->N && bd != 2 * ctx->N) throw bad_wire("ciphertext_b_dim must be d or 2d");
    if (r->num_column_ciphertexts && !r->column_ciphertexts) throw bad_wire("column_ciphertexts buffer is null");
}

template <typename Emit>
void emit_response(const WireParams* ctx, const inspire_response* r, Emit&& ct) {
    const size_t KN = (size_t)ctx->K * ctx->N;
    const uint32_t bd = r->ciphertext_b_dim ? r->ciphertext_b_dim : ctx->N;
    ct(r->ciphertext, r->ciphertext + KN, bd, -1);
    for (uint32_t c = 0; c < r->num_column_ciphertexts; c++) {
        const uint64_t* p = r->column_ciphertexts + (size_t)c * 2 * KN;
        ct(p, p + KN, ctx->N, (int)c);
    }
}

void write_response_bin(const WireParams* ctx, const inspire_response* r, Sink& s) {
    Bin w{s};
    emit_response(ctx, r, [&](const uint64_t* a, const uint64_t* b, uint32_t bd, int col) {
        bin_poly(w, ctx, a, ctx->N);
        bin_poly(w, ctx, b, bd);
        if (col < 0) w.u64(r->num_column_ciphertexts);  // Vec<RlweCiphertext> length follows the ciphertext
    });
}

void write_response_json(const WireParams* ctx, const inspire_response* r, Sink& s) {
    Json j{s};
    j.raw("{\"ciphertext\":");
    emit_response(ctx, r, [&](const uint64_t* a, const uint64_t* b, uint32_t bd, int col) {
        if (col == 0) j.raw(",\"column_ciphertexts\":[");
        else if (col > 0) j.raw(",");
        j.raw("{\"a\":");
        json_poly(j, ctx, a, ctx->N);
        j.raw(",\"b\":");
        json_poly(j, ctx, b, bd);
        j.raw("}");
    });
    if (r->num_column_ciphertexts == 0) j.raw(",\"column_ciphertexts\":[");
    j.raw("]}");
}

template <typename Write>
int to_buffer(const WireParams* ctx, const inspire_response* r, uint8_t* out, uint64_t capacity, uint64_t* len,
              Write write) {
    check_response(ctx, r);
    BufSink count;
    write(ctx, r, count);
    if (len) *len = count.n;
    if (!out || capacity < count.n)
        throw Error{INSPIRE_ERR_CAPACITY, "output buffer too small: the encoding takes " + std::to_string(count.n) + " bytes"};
    BufSink fill;
    fill.out = out;
    write(ctx, r, fill);
    return INSPIRE_OK;
}

// ---------------------------------------------------------------------------
// bincode reader
// ---------------------------------------------------------------------------

struct BinReader {
    const uint8_t* p;
    size_t n, i = 0;
    void need(size_t k) {
        if (n - i < k) throw bad_wire("bincode: unexpected end of input");
    }
    uint8_t u8() { need(1); return p[i++]; }
    uint32_t u32() { need(4); uint32_t v; std::memcpy(&v, p + i, 4); i += 4; return v; }
    uint64_t u64() { need(8); uint64_t v; std::memcpy(&v, p + i, 8); i += 8; return v; }
    size_t left() const { return n - i; }
};

// Checks a Poly's metadata against the context; appends its residue words.
void check_poly_meta(const WireParams* ctx, uint64_t nmod, const uint64_t* moduli, uint64_t q, uint64_t dim,
                     uint64_t ncoeffs, bool is_ntt, uint32_t want_dim) {
    if (nmod != ctx->K) throw bad_wire("polynomial moduli differ from the parameters'");
    for (uint32_t k = 0; k < ctx->K; k++)
        if (moduli[k] != ctx->moduli[k]) throw bad_wire("polynomial moduli differ from the parameters'");
    if (q != ctx->q) throw bad_wire("polynomial modulus q differs from the parameters'");
    if (dim != want_dim || ncoeffs != (uint64_t)ctx->K * want_dim)
        throw bad_wire("polynomial dimension " + std::to_string(dim) + " differs from the expected " +
                       std::to_string(want_dim));
    if (is_ntt) throw bad_wire("polynomials must be in the coefficient domain");
}

void check_residues(const WireParams* ctx, const uint64_t* w, uint32_t dim) {
    for (uint32_t k = 0; k < ctx->K; k++)
        for (uint32_t i = 0; i < dim; i++)
            if (w[(size_t)k * dim + i] >= ctx->moduli[k]) throw bad_wire("a polynomial residue is not reduced");
}

void bin_read_poly(BinReader& r, const WireParams* ctx, uint64_t* dst, uint32_t want_dim) {
    const uint64_t n = r.u64();
    if (n != (uint64_t)ctx->K * want_dim) throw bad_wire("polynomial length differs from K x d");
    r.need(n * 8);
    std::memcpy(dst, r.p + r.i, n * 8);
    r.i += n * 8;
    const uint64_t nmod = r.u64();
    if (nmod > 2) throw bad_wire("polynomial moduli differ from the parameters'");
    uint64_t m[2] = {0, 0};
    for (uint64_t k = 0; k < nmod; k++) m[k] = r.u64();
    const uint64_t q = r.u64(), dim = r.u64();
    (void)r.u64();  // crt_q0_inv_mod_q1: derived from the moduli
    const uint8_t ntt = r.u8();
    check_poly_meta(ctx, nmod, m, q, dim, n, ntt != 0, want_dim);
    check_residues(ctx, dst, want_dim);
}

void check_gadget(const WireParams* ctx, uint64_t base, uint64_t len) {
    if (base != ctx->gadget_base || len != ctx->ell)
        throw Error{INSPIRE_ERR_UNSUPPORTED, "the query's gadget differs from the parameters' (base " +
                                                 std::to_string(base) + ", len " + std::to_string(len) + ")"};
}

// ---------------------------------------------------------------------------
// JSON reader (the subset serde_json writes for these types)
// ---------------------------------------------------------------------------

struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    uint64_t num = 0;
    std::string str;
    std::vector<uint64_t> nums;  // an array of unsigned integers only
    bool num_array = false;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const char* key) const {
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* e;
    int depth = 0;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
    }
    [[noreturn]] void fail(const char* what) { throw bad_wire(std::string("json: ") + what); }
    bool lit(const char* t) {
        const size_t k = std::strlen(t);
        if ((size_t)(e - p) >= k && std::memcmp(p, t, k) == 0) { p += k; return true; }
        return false;
    }
    uint64_t number() {
        if (p < e && *p == '-') fail("negative numbers are not valid here");
        uint64_t v = 0;
        auto r = std::from_chars(p, e, v);
        if (r.ec != std::errc()) fail("expected an unsigned integer");
        p = r.ptr;
        if (p < e && (*p == '.' || *p == 'e' || *p == 'E')) fail("expected an integer");
        return v;
    }
    std::string string() {
        if (p >= e || *p != '"') fail("expected a string");
        p++;
        std::string s;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) fail("bad escape");
                const char c = *p++;
                switch (c) {
                    case '"': case '\\': case '/': s += c; break;
                    case 'n': s += '\n'; break;
                    case 't': s += '\t'; break;
                    case 'r': s += '\r'; break;
                    case 'b': s += '\b'; break;
                    case 'f': s += '\f'; break;
                    default: fail("unsupported escape");
                }
            } else {
                s += *p++;
            }
        }
        if (p >= e) fail("unterminated string");
        p++;
        return s;
    }
    void value(JVal& v) {
        if (++depth > 64) fail("nesting too deep");
        ws();
        if (p >= e) fail("unexpected end of input");
        const char c = *p;
        if (c == '{') {
            v.kind = JVal::OBJ;
            p++;
            ws();
            if (p < e && *p == '}') { p++; depth--; return; }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (p >= e || *p != ':') fail("expected ':'");
                p++;
                v.obj.emplace_back(std::move(k), JVal{});
                value(v.obj.back().second);
                ws();
                if (p < e && *p == ',') { p++; continue; }
                if (p < e && *p == '}') { p++; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JVal::ARR;
            v.num_array = true;
            p++;
            ws();
            if (p < e && *p == ']') { p++; depth--; return; }
            for (;;) {
                ws();
                if (v.num_array && p < e && *p >= '0' && *p <= '9') {
                    v.nums.push_back(number());
                } else {
                    if (v.num_array && !v.nums.empty()) fail("mixed array");
                    v.num_array = false;
                    v.arr.emplace_back();
                    value(v.arr.back());
                }
                ws();
                if (p < e && *p == ',') { p++; continue; }
                if (p < e && *p == ']') { p++; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JVal::BOOL;
        } else if (lit("null")) {
            v.kind = JVal::NUL;
        } else {
            v.kind = JVal::NUM;
            v.num = number();
        }
        depth--;
    }
};

const JVal& field(const JVal& o, const char* key) {
    if (o.kind != JVal::OBJ) throw bad_wire(std::string("json: expected an object holding '") + key + "'");
    const JVal* v = o.get(key);
    if (!v) throw bad_wire(std::string("json: missing field '") + key + "'");
    return *v;
}

uint64_t jnum(const JVal& v, const char* what) {
    if (v.kind != JVal::NUM) throw bad_wire(std::string("json: '") + what + "' must be an unsigned integer");
    return v.num;
}

// A JSON array of numbers (possibly empty).
const std::vector<uint64_t>& jnums(const JVal& v, const char* what) {
    static const std::vector<uint64_t> empty;
    if (v.kind != JVal::ARR || (!v.num_array && !v.arr.empty()))
        throw bad_wire(std::string("json: '") + what + "' must be an array of integers");
    return v.num_array ? v.nums : empty;
}

const std::vector<JVal>& jarr(const JVal& v, const char* what) {
    static const std::vector<JVal> empty;
    if (v.kind != JVal::ARR || (v.num_array && !v.nums.empty()))
        throw bad_wire(std::string("json: '") + what + "' must be an array");
    return v.num_array ? empty : v.arr;
}

void json_read_poly(const JVal& v, const WireParams* ctx, uint64_t* dst) {
    const auto& c = jnums(field(v, "coeffs"), "coeffs");
    const auto& m = jnums(field(v, "moduli"), "moduli");
    if (m.size() > 2) throw bad_wire("polynomial moduli differ from the parameters'");
    const JVal* ntt = v.get("is_ntt");
    const bool is_ntt = ntt && ntt->kind == JVal::BOOL && ntt->b;
    check_poly_meta(ctx, m.size(), m.data(), jnum(field(v, "q"), "q"), jnum(field(v, "dim"), "dim"), c.size(), is_ntt,
                    ctx->N);
    std::memcpy(dst, c.data(), c.size() * 8);
    check_residues(ctx, dst, ctx->N);
}

void finish_query(inspire_wire_query& w, inspire_query* q, uint32_t seeded) {
    q->db_id = 0;
    q->seeded = seeded;
    q->rgsw = w.rgsw.data();
    q->rgsw_seeds = seeded ? w.seeds.data() : nullptr;
    q->y_all_len = 0;
    q->y_all = nullptr;
    q->y_body = w.y_body.empty() ? nullptr : w.y_body.data();
}

}  // namespace
}  // namespace inspire

extern "C" {

int inspire_response_to_binary(const inspire_params* params, const inspire_response* r, uint8_t* out,
                               uint64_t capacity, uint64_t* len) {
    INSPIRE_TRY({
        const WireParams wp = wire_params(params);
        return to_buffer(&wp, r, out, capacity, len, write_response_bin);
    })
}

int inspire_response_to_json(const inspire_params* params, const inspire_response* r, char* out, uint64_t capacity,
                             uint64_t* len) {
    INSPIRE_TRY({
        const WireParams wp = wire_params(params);
        return to_buffer(&wp, r, reinterpret_cast<uint8_t*>(out), capacity, len, write_response_json);
    })
}

int inspire_response_from_binary(const inspire_params* params, const uint8_t* in, uint64_t len,
                                 inspire_response* r) {
    INSPIRE_TRY({
        const WireParams wp = wire_params(params);
        const WireParams* ctx = &wp;
        if (!in || !r || !r->ciphertext) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        const size_t KN = (size_t)ctx->K * ctx->N;
        BinReader rd{in, (size_t)len};
        // the b dimension of the ciphertext: read ahead to its coefficient count
        BinReader peek = rd;
        const uint64_t na = peek.u64();
        if (na != KN) throw bad_wire("ciphertext a: polynomial length differs from K x d");
        peek.need(na * 8 + 8);
        peek.i += na * 8;
        const uint64_t nm = peek.u64();
        if (nm > 2) throw bad_wire("polynomial moduli differ from the parameters'");
        peek.need(nm * 8 + 8 * 3 + 1);
        peek.i += nm * 8 + 8 * 3 + 1;
        const uint64_t nb = peek.u64();
        const uint32_t bd = (uint32_t)(nb / ctx->K);
        if (nb % ctx->K || (bd != ctx->N && bd != 2 * ctx->N)) throw bad_wire("ciphertext b: unexpected length");
        const size_t cap = r->ciphertext_capacity ? r->ciphertext_capacity : 2 * KN;
        if (cap < KN + (size_t)ctx->K * bd) throw Error{INSPIRE_ERR_CAPACITY, "ciphertext capacity too small"};
        bin_read_poly(rd, ctx, r->ciphertext, ctx->N);
        bin_read_poly(rd, ctx, r->ciphertext + KN, bd);
        r->ciphertext_b_dim = bd;
        const uint64_t cols = rd.u64();
        if (cols > 0x7fffffffull) throw bad_wire("too many column ciphertexts");
        if (cols && (!r->column_ciphertexts || r->column_capacity < cols))
            throw Error{INSPIRE_ERR_CAPACITY, "column_ciphertexts capacity too small"};
        for (uint64_t c = 0; c < cols; c++) {
            bin_read_poly(rd, ctx, r->column_ciphertexts + c * 2 * KN, ctx->N);
            bin_read_poly(rd, ctx, r->column_ciphertexts + c * 2 * KN + KN, ctx->N);
        }
        r->num_column_ciphertexts = (uint32_t)cols;
        if (rd.left()) throw bad_wire("bincode: trailing bytes after the response");
        return INSPIRE_OK;
    })
}

int inspire_query_from_binary(const inspire_params* params, const uint8_t* in, uint64_t len, uint32_t seeded,
                              inspire_wire_query** out, inspire_query* q) {
    INSPIRE_TRY({
        const WireParams wp = wire_params(params);
        const WireParams* ctx = &wp;
        if (!in || !out || !q) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *out = nullptr;
        const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
        std::unique_ptr<inspire_wire_query> w(new inspire_wire_query());
        BinReader rd{in, (size_t)len};
        std::memset(q, 0, sizeof *q);
        q->shard_id = rd.u32();
        const uint64_t rows = rd.u64();
        if (rows != 2 * ell) throw Error{INSPIRE_ERR_UNSUPPORTED, "the RGSW query must have 2 x gadget_len rows"};
        if (seeded) {
            w->rgsw.resize(rows * KN);
            w->seeds.resize(rows * 32);
            for (uint64_t i = 0; i < rows; i++) {
                rd.need(32);
                std::memcpy(&w->seeds[i * 32], rd.p + rd.i, 32);
                rd.i += 32;
                bin_read_poly(rd, ctx, &w->rgsw[i * KN], ctx->N);
            }
        } else {
            w->rgsw.resize(rows * 2 * KN);
            for (uint64_t i = 0; i < rows; i++) {
                bin_read_poly(rd, ctx, &w->rgsw[i * 2 * KN], ctx->N);
                bin_read_poly(rd, ctx, &w->rgsw[i * 2 * KN + KN], ctx->N);
            }
        }
        const uint64_t gbase = rd.u64(), glen = rd.u64();
        (void)rd.u64();  // gadget q
        check_gadget(ctx, gbase, glen);
        const uint32_t mode = rd.u32();
        if (mode > 1) throw bad_wire("unknown packing_mode variant " + std::to_string(mode));
        q->packing_mode = mode;
        // inspiring_packing_keys: left out when None (skip_serializing_if)
        if (rd.left()) {
            const uint8_t tag = rd.u8();
            if (tag > 1) throw bad_wire("bincode: bad Option tag");
            if (tag == 1) {
                q->has_packing_keys = 1;
                const uint64_t ny = rd.u64();
                if (ny && ny != ell) throw Error{INSPIRE_ERR_UNSUPPORTED, "y_body must hold gadget_len polynomials"};
                w->y_body.resize(ny * KN);
                for (uint64_t i = 0; i < ny; i++) bin_read_poly(rd, ctx, &w->y_body[i * KN], ctx->N);
                // z_body is written only when non-empty (full packing): then a
                // u64 count and that many polynomials precede full_key and
                // num_to_pack (9 bytes)
                if (rd.left() < 9) throw bad_wire("bincode: unexpected end of input");
                if (rd.left() > 9) {
                    BinReader z = rd;
                    const uint64_t nz = z.u64();
                    const uint64_t poly_bytes = 8 + (uint64_t)KN * 8 + 8 + 8 * (uint64_t)ctx->K + 8 * 3 + 1;
                    if (nz && nz <= 64 && z.left() == nz * poly_bytes + 9)
                        throw Error{INSPIRE_ERR_UNSUPPORTED, "full packing keys (z_body) are not on this path"};
                    throw bad_wire("bincode: trailing bytes after the query");
                }
                (void)rd.u8();   // full_key
                (void)rd.u64();  // num_to_pack
            }
        }
        if (rd.left()) throw bad_wire("bincode: trailing bytes after the query");
        finish_query(*w, q, seeded);
        *out = w.release();
        return INSPIRE_OK;
    })
}

int inspire_query_from_json(const inspire_params* params, const char* json, uint64_t len, inspire_wire_query** out,
                            inspire_query* q) {
    INSPIRE_TRY({
        const WireParams wp = wire_params(params);
        const WireParams* ctx = &wp;
        if (!json || !out || !q) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *out = nullptr;
        const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
        JParser jp{json, json + len};
        JVal root;
        jp.value(root);
        jp.ws();
        if (jp.p != jp.e) throw bad_wire("json: trailing characters after the query");
        std::unique_ptr<inspire_wire_query> w(new inspire_wire_query());
        std::memset(q, 0, sizeof *q);
        const uint64_t sid = jnum(field(root, "shard_id"), "shard_id");
        if (sid > 0xffffffffull) throw bad_wire("shard_id out of range");
        q->shard_id = (uint32_t)sid;
        const JVal& rg = field(root, "rgsw_ciphertext");
        const JVal& gad = field(rg, "gadget");
        check_gadget(ctx, jnum(field(gad, "base"), "base"), jnum(field(gad, "len"), "len"));
        const auto& rows = jarr(field(rg, "rows"), "rows");
        if (rows.size() != 2 * ell) throw Error{INSPIRE_ERR_UNSUPPORTED, "the RGSW query must have 2 x gadget_len rows"};
        const uint32_t seeded = rows[0].kind == JVal::OBJ && rows[0].get("seed") ? 1u : 0u;
        if (seeded) {
            w->rgsw.resize(rows.size() * KN);
            w->seeds.resize(rows.size() * 32);
            for (size_t i = 0; i < rows.size(); i++) {
                const auto& sd = jnums(field(rows[i], "seed"), "seed");
                if (sd.size() != 32) throw bad_wire("a seed must hold 32 bytes");
                for (int b = 0; b < 32; b++) {
                    if (sd[b] > 255) throw bad_wire("a seed byte is out of range");
                    w->seeds[i * 32 + b] = (uint8_t)sd[b];
                }
                json_read_poly(field(rows[i], "b"), ctx, &w->rgsw[i * KN]);
            }
        } else {
            w->rgsw.resize(rows.size() * 2 * KN);
            for (size_t i = 0; i < rows.size(); i++) {
                json_read_poly(field(rows[i], "a"), ctx, &w->rgsw[i * 2 * KN]);
                json_read_poly(field(rows[i], "b"), ctx, &w->rgsw[i * 2 * KN + KN]);
            }
        }
        q->packing_mode = INSPIRE_PACKING_INSPIRING;  // #[serde(default)]
        if (const JVal* pm = root.get("packing_mode")) {
            if (pm->kind != JVal::STR) throw bad_wire("json: packing_mode must be a string");
            if (pm->str == "tree") q->packing_mode = INSPIRE_PACKING_TREE;
            else if (pm->str != "inspiring") throw bad_wire("unknown packing_mode '" + pm->str + "'");
        }
        const JVal* keys = root.get("inspiring_packing_keys");
        if (keys && keys->kind != JVal::NUL) {
            q->has_packing_keys = 1;
            const auto& yb = jarr(field(*keys, "y_body"), "y_body");
            if (!yb.empty() && yb.size() != ell) throw Error{INSPIRE_ERR_UNSUPPORTED, "y_body must hold gadget_len polynomials"};
            if (const JVal* zb = keys->get("z_body"))
                if (!jarr(*zb, "z_body").empty()) throw Error{INSPIRE_ERR_UNSUPPORTED, "full packing keys (z_body) are not on this path"};
            w->y_body.resize(yb.size() * KN);
            for (size_t i = 0; i < yb.size(); i++) json_read_poly(yb[i], ctx, &w->y_body[i * KN]);
        }
        finish_query(*w, q, seeded);
        *out = w.release();
        return INSPIRE_OK;
    })
}

void inspire_wire_query_free(inspire_wire_query* w) { delete w; }

}  // extern "C"
