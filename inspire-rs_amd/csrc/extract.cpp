// Client-side extract (src/pir/extract.rs) over the device: host driver and
// the C ABI entry points inspire_extract*, inspire_extract_batch_device.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "abi.hpp"
#include "context.hpp"
#include "extract.hpp"
#include "modarith.hpp"
#include "respond.hpp"

namespace inspire {

// mod_inverse(d, p).unwrap_or(1) (extract.rs:141, 196-213)
static uint64_t packed_multiplier(uint64_t d, uint64_t p) {
    if (p == 1) return 0;
    // extended Euclid on (d mod p, p)
    __int128 r0 = (__int128)p, r1 = (__int128)(d % p), t0 = 0, t1 = 1;
    while (r1) {
        const __int128 qq = r0 / r1, r2 = r0 - qq * r1, t2 = t0 - qq * t1;
        r0 = r1; r1 = r2; t0 = t1; t1 = t2;
    }
    if (r0 != 1) return 1;
    t0 %= (__int128)p;
    if (t0 < 0) t0 += p;
    return (uint64_t)t0;
}

static ExtractArgs make_args(const inspire_ctx* ctx, const inspire_extract_batch& b) {
    const inspire_ctx* R = root(ctx);
    if (b.mode > INSPIRE_EXTRACT_INSPIRING)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "extract mode must be INSPIRE_EXTRACT_COLUMNS, _PACKED or _INSPIRING"};
    const uint32_t N = ctx->N, K = ctx->K;
    const uint32_t b_dim = b.b_dim ? b.b_dim : N;
    if (b_dim != N && b_dim != 2 * N) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "b_dim must be d or 2d"};
    ExtractArgs A{};
    A.count = b.count;
    A.mode = b.mode;
    A.entry_size = b.entry_size;
    A.nvals = (uint32_t)(((uint64_t)b.entry_size * 8 + 15) / 16);  // num_columns (extract.rs:42)
    const uint32_t ncols = b.mode == INSPIRE_EXTRACT_COLUMNS ? b.num_column_ciphertexts : 0;
    if (b.mode != INSPIRE_EXTRACT_COLUMNS && A.nvals > N)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT,
                    "entry_size " + std::to_string(b.entry_size) + " needs coefficients beyond d = " +
                        std::to_string(N) + " of the packed ciphertext"};
    A.use_cols = ncols ? 1u : 0u;
    A.ndec = b.mode != INSPIRE_EXTRACT_COLUMNS ? A.nvals : (ncols ? umin32(ncols, A.nvals) : 1u);
    A.b_off = K * N;
    A.col_off = K * N + K * b_dim;
    const uint64_t min_stride = (uint64_t)A.col_off + (uint64_t)b.num_column_ciphertexts * 2 * K * N;
    A.stride = b.stride ? b.stride : min_stride;
    if (A.stride < (b.mode == INSPIRE_EXTRACT_COLUMNS ? min_stride : (uint64_t)A.col_off))
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "stride is smaller than one response"};
    A.delta = R->delta;
    A.half_delta = R->delta / 2;
    A.p = R->params.p;
    A.d_inv = packed_multiplier(N, R->params.p);
    for (uint32_t k = 0; k < K; k++) A.r64[k] = (uint32_t)pow2_64_mod(ctx->hm[k].q);
    return A;
}

void extract_batch(inspire_ctx* ctx, const inspire_extract_batch& b) {
    if (!b.count) return;
    if (!b.responses || !b.secret_keys || (!b.entries && b.entry_size))
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null device pointer"};
    ExtractArgs A = make_args(ctx, b);
    A.resp = b.responses;
    A.sk = b.secret_keys;
    A.entries = b.entries;
    {
        prof_launch(ctx, P_EXTRACT, "extract", [A, dev = ctx->dev, st = ctx->stream](LaunchEv ev) {
            return launch_extract(A, dev, st, ev.start, ev.stop);
        });
    }
    hip_check(hipStreamSynchronize(ctx->stream), "extract sync");
}

// One response from host buffers: the ciphertext, the column ciphertexts the
// mode reads, and the key go up as u32 words; the entry comes back.
static void extract_host(inspire_ctx* ctx, const inspire_response* r, const uint64_t* sk, uint32_t entry_size,
                         uint32_t mode, uint8_t* entry) {
    if (!r || !sk || (!entry && entry_size) || !r->ciphertext)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
    const uint32_t N = ctx->N, K = ctx->K;
    const uint32_t b_dim = r->ciphertext_b_dim ? r->ciphertext_b_dim : N;
    if (b_dim != N && b_dim != 2 * N) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "ciphertext_b_dim must be d or 2d"};
    if (r->num_column_ciphertexts && !r->column_ciphertexts)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "column_ciphertexts is null"};
    const uint32_t nvals = (uint32_t)(((uint64_t)entry_size * 8 + 15) / 16);
    const uint32_t ncols =
        mode == INSPIRE_EXTRACT_COLUMNS ? umin32(r->num_column_ciphertexts, nvals) : 0u;  // the ones read
    const size_t KN = (size_t)K * N, ct_words = KN + (size_t)K * b_dim;
    const size_t words = ct_words + (size_t)ncols * 2 * KN;
    uint64_t qmax = 0;
    for (uint32_t k = 0; k < K; k++) qmax = ctx->hm[k].q > qmax ? ctx->hm[k].q : qmax;
    auto word = [&](size_t i) -> uint64_t {
        return i < ct_words ? r->ciphertext[i] : r->column_ciphertexts[i - ct_words];
    };
    std::vector<uint32_t> h(words + KN);
    for (size_t i = 0; i < words; i++) {
        const uint64_t v = word(i);
        // a and every column residue belong to residue (i / N) mod K; the
        // ciphertext's b words only need to fit the widest modulus
        const bool is_b = i >= KN && i < ct_words;
        const uint64_t bound = is_b ? qmax : ctx->hm[((i < ct_words ? i : i - ct_words) / N) % K].q;
        if (v >= bound) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "response residue out of range"};
        h[i] = (uint32_t)v;
    }
    for (size_t i = 0; i < KN; i++) {
        if (sk[i] >= ctx->hm[i / N].q) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "secret key residue out of range"};
        h[words + i] = (uint32_t)sk[i];
    }
    ctx->ws0.reserve(h.size() * 4);
    ctx->ws1.reserve(entry_size ? entry_size : 1);
    uint32_t* d = ctx->ws0.get<uint32_t>();
    hip_check(hipMemcpyAsync(d, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream), "extract upload");
    inspire_extract_batch b{};
    b.count = 1;
    b.mode = mode;
    b.entry_size = entry_size;
    b.num_column_ciphertexts = ncols;
    b.b_dim = b_dim;
    b.stride = words;
    b.responses = d;
    b.secret_keys = d + words;
    b.entries = ctx->ws1.get<uint8_t>();
    extract_batch(ctx, b);
    if (entry_size)
        hip_check(hipMemcpyAsync(entry, b.entries, entry_size, hipMemcpyDeviceToHost, ctx->stream), "extract download");
    hip_check(hipStreamSynchronize(ctx->stream), "extract sync");
}

}  // namespace inspire

using inspire::Error;

extern "C" {

int inspire_extract(inspire_ctx* ctx, const inspire_response* r, const uint64_t* secret_key, uint32_t entry_size,
                    uint8_t* entry) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        inspire::extract_host(ctx, r, secret_key, entry_size, INSPIRE_EXTRACT_COLUMNS, entry);
        return INSPIRE_OK;
    })
}

int inspire_extract_with_variant(inspire_ctx* ctx, const inspire_response* r, const uint64_t* secret_key,
                                 uint32_t entry_size, uint32_t variant, uint8_t* entry) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (variant > INSPIRE_VARIANT_TWO_PACKING) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "unknown variant"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        inspire::extract_host(ctx, r, secret_key, entry_size,
                              variant == INSPIRE_VARIANT_NO_PACKING ? INSPIRE_EXTRACT_COLUMNS : INSPIRE_EXTRACT_PACKED,
                              entry);
        return INSPIRE_OK;
    })
}

int inspire_extract_inspiring(inspire_ctx* ctx, const inspire_response* r, const uint64_t* secret_key,
                              uint32_t entry_size, uint8_t* entry) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        inspire::extract_host(ctx, r, secret_key, entry_size, INSPIRE_EXTRACT_INSPIRING, entry);
        return INSPIRE_OK;
    })
}

int inspire_extract_batch_device(inspire_ctx* ctx, const inspire_extract_batch* batch) {
    INSPIRE_TRY({
        if (!ctx || !batch) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        inspire::extract_batch(ctx, *batch);
        return INSPIRE_OK;
    })
}

}  // extern "C"
