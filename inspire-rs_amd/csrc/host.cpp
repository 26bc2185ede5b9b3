// Host side of the MI355X respond path: the C ABI of include/inspire_mi355x.h.
//
// The context owns every device table; respond calls stage query data into
// reusable device workspaces, run the kernels on one HIP stream, and copy the
// response back.  Batches are grouped by database and variant so that one
// launch of each kernel serves every query of a group.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "modarith.hpp"
#include "abi.hpp"
#include "respond.hpp"
#include "tree.hpp"
#include "ntt_mfma.hpp"

using namespace inspire;

namespace inspire {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void DevBuf::alloc(size_t bytes) {
    p_ = nullptr;
    n_ = 0;
    if (!bytes) return;
    if (hipMalloc(&p_, bytes) != hipSuccess) {
        p_ = nullptr;
        (void)hipGetLastError();
        throw Error{INSPIRE_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed"};
    }
    n_ = bytes;
}

void DevBuf::release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
}

void PinBuf::reserve(size_t bytes) {
    if (bytes <= n_) return;
    release();
    if (hipHostMalloc(&p_, bytes, hipHostMallocDefault) != hipSuccess) {
        p_ = nullptr;
        (void)hipGetLastError();
        throw Error{INSPIRE_ERR_OUT_OF_MEMORY, "pinned host allocation of " + std::to_string(bytes) + " bytes failed"};
    }
    n_ = bytes;
}

void PinBuf::release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error{INSPIRE_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

// ---------------------------------------------------------------------------
// Parameter set and tables
// ---------------------------------------------------------------------------

static uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

// find_primitive_root (src/math/ntt.rs:525-538): the first g >= 2 whose
// g^((q-1)/2n) has order exactly 2n.
static uint64_t reference_psi(uint64_t q, uint64_t two_n) {
    const uint64_t e = (q - 1) / two_n;
    for (uint64_t g = 2; g < q; g++) {
        uint64_t c = powmod64(g, e, q);
        if (powmod64(c, two_n, q) == 1 && powmod64(c, two_n / 2, q) != 1) return c;
    }
    throw Error{INSPIRE_ERR_INVALID_PARAMS, "no primitive 2n-th root of unity"};
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) { uint64_t t = a % b; a = b; b = t; }
    return a;
}

// InspireParams::validate (src/params.rs:348-387), same messages.
static void validate_params(const inspire_params& p) {
    auto bad = [](const char* m) { throw Error{INSPIRE_ERR_INVALID_PARAMS, m}; };
    if (p.ring_dim == 0 || (p.ring_dim & (p.ring_dim - 1))) bad("ring_dim must be a power of two");
    if (p.crt_count == 0) bad("crt_moduli must be non-empty");
    if (p.crt_count > 2) bad("crt_moduli length > 2 is not supported");
    const uint64_t two_n = 2ull * p.ring_dim;
    for (uint32_t k = 0; k < p.crt_count; k++)
        if (p.crt_moduli[k] % two_n != 1) bad("CRT moduli must be ≡ 1 (mod 2d) for NTT");
    unsigned __int128 prod = 1;
    for (uint32_t k = 0; k < p.crt_count; k++) prod *= p.crt_moduli[k];
    if (prod != (unsigned __int128)p.q) bad("q must equal product of CRT moduli");
    if (p.crt_count == 2 && gcd64(p.crt_moduli[0], p.crt_moduli[1]) != 1) bad("CRT moduli must be coprime");
    if (p.q < p.p) bad("q must be >= p");
}

// Shape limits of this implementation (every other parameter set the
// reference accepts is refused with INSPIRE_ERR_UNSUPPORTED).
static void check_supported(const inspire_params& p) {
    auto no = [](const std::string& m) { throw Error{INSPIRE_ERR_UNSUPPORTED, m}; };
    if (p.ring_dim < 256 || p.ring_dim > 4096) no("ring_dim must be in [256, 4096] on the GPU path");
    for (uint32_t k = 0; k < p.crt_count; k++)
        if (p.crt_moduli[k] >= (1ull << 29))
            no("each CRT modulus must be below 2^29 on the GPU path (32-bit residue arithmetic)");
    if (p.gadget_base < 2 || (p.gadget_base & (p.gadget_base - 1))) no("gadget_base must be a power of two");
    if (p.gadget_base > (1ull << 32)) no("gadget_base must be at most 2^32 on the GPU path");
    if (p.gadget_len == 0 || p.gadget_len > (uint32_t)kMaxEll) no("gadget_len must be in [1, 4]");
    if (p.p == 0) no("p must be positive");
}

static void build_modulus(HostModulus& m, uint64_t q, uint32_t n, uint32_t logn) {
    m.q = q;
    m.psi = reference_psi(q, 2ull * n);
    const uint64_t psi_inv = powmod64(m.psi, q - 2, q);
    m.tw.assign(n, 0);
    m.twp.assign(n, 0);
    m.itw.assign(n, 0);
    m.itwp.assign(n, 0);
    for (uint32_t i = 1; i < n; i++) {  // tw[i] = psi^br(i) (ntt.rs:541-583, without the 2^64 factor)
        const uint32_t e = bitrev(i, logn);
        m.tw[i] = (uint32_t)powmod64(m.psi, e, q);
        m.itw[i] = (uint32_t)powmod64(psi_inv, e, q);
        m.twp[i] = shoup_pre(m.tw[i], (uint32_t)q);
        m.itwp[i] = shoup_pre(m.itw[i], (uint32_t)q);
    }
}

inspire_ctx* make_lane_ctx(inspire_ctx* owner) {
    std::unique_ptr<inspire_ctx> c(new inspire_ctx());
    c->params = owner->params;
    c->N = owner->N;
    c->K = owner->K;
    c->logn = owner->logn;
    c->ell = owner->ell;
    c->q = owner->q;
    c->delta = owner->delta;
    std::memcpy(c->w_seed, owner->w_seed, 32);
    for (uint32_t k = 0; k < owner->K; k++) c->hm[k] = owner->hm[k];
    c->dev = owner->dev;  // device tables belong to the owner
    c->device = owner->device;
    c->owner = owner;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "lane stream");
    hip_check(hipEventCreateWithFlags(&c->chain_done, hipEventDisableTiming), "chain event");
    return c.release();
}

void destroy_lane_ctx(inspire_ctx* c) { inspire_ctx_destroy(c); }

namespace {

// A lane reads its owner's databases, tables and keys: it holds the owner's
// reader lock for the whole call (loads and registrations hold it exclusively).
void owner_only(const inspire_ctx* ctx, const char* what) {
    if (ctx->owner)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, std::string(what) + " goes through the lane's owner context"};
}

}  // namespace

}  // namespace inspire

// ---------------------------------------------------------------------------
// Scaling constants
// ---------------------------------------------------------------------------
namespace inspire {

ScaleDev scale_none() {
    ScaleDev s{};
    return s;
}

// Forward post-scale by c_k (c given per residue, reduced mod q_k).
ScaleDev scale_fwd(const inspire_ctx* ctx, const uint64_t* c) {
    ScaleDev s{};
    s.enable = 1;
    for (uint32_t k = 0; k < ctx->K; k++) {
        const uint32_t q = (uint32_t)ctx->hm[k].q;
        s.c[k] = (uint32_t)(c[k] % q);
        s.cp[k] = shoup_pre(s.c[k], q);
    }
    return s;
}

// Inverse with the output multiplied by c_k: last stage constants
// (n^-1 c, n^-1 c psi^-br(1)).
ScaleDev scale_inv(const inspire_ctx* ctx, const uint64_t* c) {
    ScaleDev s{};
    s.enable = 1;
    for (uint32_t k = 0; k < ctx->K; k++) {
        const uint64_t q = ctx->hm[k].q;
        const uint64_t ninv = powmod64(ctx->N, q - 2, q);
        const uint64_t cc = mulmod64(ninv, c ? c[k] % q : 1, q);
        s.c[k] = (uint32_t)cc;
        s.cp[k] = shoup_pre(s.c[k], (uint32_t)q);
        s.cw[k] = (uint32_t)mulmod64(cc, ctx->hm[k].itw[1], q);
        s.cwp[k] = shoup_pre(s.cw[k], (uint32_t)q);
    }
    return s;
}

uint64_t pow2_64_mod(uint64_t q) { return (uint64_t)(((unsigned __int128)1 << 64) % q); }

}  // namespace inspire

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* inspire_last_error(void) { return inspire::g_last_error.c_str(); }

int inspire_ctx_create(const inspire_params* params, const uint8_t w_seed[32], inspire_ctx** out) {
    INSPIRE_TRY({
        if (!params || !w_seed || !out) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *out = nullptr;
        validate_params(*params);
        check_supported(*params);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            (void)hipGetLastError();
            throw Error{INSPIRE_ERR_DEVICE, "no HIP device available"};
        }
        std::unique_ptr<inspire_ctx> ctx(new inspire_ctx());
        hip_check(hipGetDevice(&ctx->device), "current device");
        noalloc_scratch_setup(ctx->device);
        ctx->params = *params;
        ctx->N = params->ring_dim;
        ctx->K = params->crt_count;
        ctx->ell = params->gadget_len;
        ctx->q = params->q;
        ctx->delta = params->q / params->p;  // params.rs:296-298
        uint32_t logn = 0;
        while ((1u << logn) < ctx->N) logn++;
        ctx->logn = logn;
        std::memcpy(ctx->w_seed, w_seed, 32);
        for (uint32_t k = 0; k < ctx->K; k++) build_modulus(ctx->hm[k], params->crt_moduli[k], ctx->N, logn);
        // device twiddle tables: per modulus tw, twp, itw, itwp, twm
        const size_t per = (size_t)5 * ctx->N;
        ctx->tables.alloc(per * ctx->K * sizeof(uint32_t));
        std::vector<uint32_t> h(per * ctx->K);
        for (uint32_t k = 0; k < ctx->K; k++) {
            uint32_t* b = h.data() + per * k;
            std::copy(ctx->hm[k].tw.begin(), ctx->hm[k].tw.end(), b);
            std::copy(ctx->hm[k].twp.begin(), ctx->hm[k].twp.end(), b + ctx->N);
            std::copy(ctx->hm[k].itw.begin(), ctx->hm[k].itw.end(), b + 2 * ctx->N);
            std::copy(ctx->hm[k].itwp.begin(), ctx->hm[k].itwp.end(), b + 3 * ctx->N);
            const uint64_t q = ctx->hm[k].q, r32 = (1ull << 32) % q;
            for (uint32_t i = 0; i < ctx->N; i++) b[4 * ctx->N + i] = (uint32_t)mulmod64(ctx->hm[k].tw[i], r32, q);
        }
        hip_check(hipMemcpy(ctx->tables.get(), h.data(), h.size() * 4, hipMemcpyHostToDevice), "upload tables");
        CtxDev& D = ctx->dev;
        D.K = ctx->K;
        D.logn = logn;
        D.n = ctx->N;
        D.ell = ctx->ell;
        uint32_t glog = 0;
        while ((1ull << glog) < params->gadget_base) glog++;
        D.gadget_log = glog;
        D.gadget_mask = (uint32_t)(params->gadget_base - 1);
        D.q = ctx->q;
        D.small_digits = 1;
        for (uint32_t k = 0; k < ctx->K; k++)
            if (params->gadget_base > ctx->hm[k].q) D.small_digits = 0;
        for (uint32_t k = 0; k < ctx->K; k++) {
            const uint32_t q = (uint32_t)ctx->hm[k].q;
            const uint32_t* base = ctx->tables.get<uint32_t>() + per * k;
            D.m[k].q = q;
            D.m[k].q2 = 2 * q;
            D.m[k].qinv = neg_inv32(q);
            D.m[k].tw = base;
            D.m[k].twp = base + ctx->N;
            D.m[k].itw = base + 2 * ctx->N;
            D.m[k].itwp = base + 3 * ctx->N;
            D.m[k].twm = base + 4 * ctx->N;
            D.delta[k] = (uint32_t)(ctx->delta % q);
            D.deltap[k] = shoup_pre(D.delta[k], q);
            D.onep[k] = shoup_pre(1, q);
            D.mont[k] = (uint32_t)((1ull << 32) % q);
            D.montp[k] = shoup_pre(D.mont[k], q);
        }
        if (ctx->K == 2) {
            const uint32_t q1 = (uint32_t)ctx->hm[1].q;
            D.q0_inv_q1 = (uint32_t)powmod64(ctx->hm[0].q % q1, q1 - 2, q1);
            D.q0_inv_q1p = shoup_pre(D.q0_inv_q1, q1);
            D.one_q1p = shoup_pre(1, q1);
        }
        if (ctx->N == 2048) {  // the matrix-core forward transform's tables (chosen by tuning)
            bool small = true;
            uint64_t qs[kMaxK], psis[kMaxK];
            for (uint32_t k = 0; k < ctx->K; k++) {
                qs[k] = ctx->hm[k].q;
                psis[k] = ctx->hm[k].psi;
                small = small && qs[k] < (1ull << 28);
            }
            if (small) {
                const MfmaNttHost H = ntt_mfma_tables(ctx->K, qs, psis);
                ctx->mfma_tables.alloc(H.bytes.size());
                hip_check(hipMemcpy(ctx->mfma_tables.get(), H.bytes.data(), H.bytes.size(), hipMemcpyHostToDevice),
                          "mfma tables");
                const uint8_t* b = ctx->mfma_tables.get<uint8_t>();
                D.mfma.tw1 = reinterpret_cast<const uint32_t*>(b + H.o_tw1);
                D.mfma.tw2 = reinterpret_cast<const uint32_t*>(b + H.o_tw2);
                D.mfma.w64 = reinterpret_cast<const int8_t*>(b + H.o_w64);
                D.mfma.w32 = reinterpret_cast<const int8_t*>(b + H.o_w32);
                D.mfma.c1 = reinterpret_cast<const uint32_t*>(b + H.o_c1);
            }
        }
        hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream create");
        hip_check(hipEventCreateWithFlags(&ctx->chain_done, hipEventDisableTiming), "chain event");
        *out = ctx.release();
        return INSPIRE_OK;
    })
}

void inspire_ctx_destroy(inspire_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    graph_cache_clear(ctx->graphs);
    for (auto& s : ctx->prof.spans) {
        (void)hipEventDestroy(s.a);
        (void)hipEventDestroy(s.b);
    }
    for (auto e : ctx->prof.pool) (void)hipEventDestroy(e);
    if (ctx->chain_done) {
        inspire_ctx* R = root(ctx);
        {
            std::lock_guard<std::mutex> lk(R->chain_mu);
            if (R->last_chain == ctx->chain_done) R->last_chain = nullptr;
        }
        (void)hipEventDestroy(ctx->chain_done);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

uint64_t inspire_ctx_device_bytes(const inspire_ctx* ctx) {
    if (!ctx) return 0;
    uint64_t b = ctx->tables.bytes() + ctx->clients.keys.bytes() + ctx->ws0.bytes() + ctx->ws1.bytes() + ctx->ws2.bytes() + ctx->ws3.bytes() +
                 ctx->ws4.bytes() + ctx->ws5.bytes() + ctx->ws6.bytes() + ctx->galois.bytes() + ctx->yhat.bytes();
    for (auto& kv : ctx->dbs) b += kv.second.digits.bytes();
    std::lock_guard<std::mutex> lk(const_cast<inspire_ctx*>(ctx)->packs_mu);
    for (auto& kv : ctx->packs) b += kv.second->stepu.bytes() + kv.second->perm.bytes() + kv.second->w_all.bytes();
    return b;
}

int inspire_ctx_create_lane(inspire_ctx* ctx, inspire_ctx** out) {
    INSPIRE_TRY({
        if (!ctx || !out) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *out = nullptr;
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "a lane cannot own lanes"};
        *out = make_lane_ctx(ctx);
        return INSPIRE_OK;
    })
}

int inspire_respond_batch_device(inspire_ctx* ctx, const inspire_device_batch* batch) {
    INSPIRE_TRY({
        if (!ctx || !batch) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        LaneRead lr(ctx);
        respond_batch_device(ctx, *batch);
        return INSPIRE_OK;
    })
}

int inspire_ctx_set_graphs(inspire_ctx* ctx, int enable) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        graph_cache_clear(ctx->graphs);
        ctx->graphs.enabled = enable != 0;
        return INSPIRE_OK;
    })
}

int inspire_ctx_set_chain_window(inspire_ctx* ctx, uint32_t window_us) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "the chain window is set on the owner context"};
        std::lock_guard<std::mutex> lk(ctx->chain_mu);
        ctx->chain_window_us = window_us;
        ctx->last_chain = nullptr;
        return INSPIRE_OK;
    })
}

int inspire_ctx_chain_window(const inspire_ctx* ctx, uint32_t* window_us) {
    INSPIRE_TRY({
        if (!ctx || !window_us) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *window_us = root(ctx)->chain_window_us.load();
        return INSPIRE_OK;
    })
}

int inspire_ctx_graph_stats(const inspire_ctx* ctx, uint64_t* captures, uint64_t* replays) {
    INSPIRE_TRY({
        if (!ctx || !captures || !replays) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *captures = ctx->graphs.captures;
        *replays = ctx->graphs.replays;
        return INSPIRE_OK;
    })
}

int inspire_ctx_set_profiling(inspire_ctx* ctx, int enable) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        prof_collect(ctx);
        ctx->prof.on = enable != 0;
        return INSPIRE_OK;
    })
}

int inspire_ctx_set_profiling_kernels(inspire_ctx* ctx, const char* kernels) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        prof_collect(ctx);
        uint32_t mask = 0;
        if (!kernels || !*kernels) {
            mask = ~0u;
        } else {
            std::string all(kernels);
            size_t pos = 0;
            while (pos <= all.size()) {
                const size_t end = std::min(all.find(',', pos), all.size());
                const std::string name = all.substr(pos, end - pos);
                int id = -1;
                for (int i = 0; i < P_COUNT; i++)
                    if (name == kProfNames[i]) id = i;
                if (id < 0) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "unknown kernel name '" + name + "'"};
                mask |= 1u << id;
                pos = end + 1;
            }
        }
        ctx->prof.mask = mask;
        return INSPIRE_OK;
    })
}

int inspire_ctx_profile(inspire_ctx* ctx, const char* kernel, double* total_ms, uint64_t* launches, int reset) {
    INSPIRE_TRY({
        if (!ctx || !kernel) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        prof_collect(ctx);
        for (int i = 0; i < P_COUNT; i++) {
            if (std::strcmp(kProfNames[i], kernel) != 0) continue;
            if (total_ms) *total_ms = ctx->prof.total_ms[i];
            if (launches) *launches = ctx->prof.launches[i];
            if (reset) {
                ctx->prof.total_ms[i] = 0;
                ctx->prof.launches[i] = 0;
            }
            return INSPIRE_OK;
        }
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, std::string("unknown kernel name ") + kernel};
    })
}

int inspire_ntt_forward(inspire_ctx* ctx, uint64_t* polys, uint64_t count) {
    INSPIRE_TRY({
        if (!ctx || (!polys && count)) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        const size_t words = (size_t)count * ctx->K * ctx->N;
        ctx->ws0.reserve(words * 4);
        upload_u64(ctx, polys, words, ctx->ws0.get<uint32_t>());
        uint64_t r64[kMaxK];
        for (uint32_t k = 0; k < ctx->K; k++) r64[k] = pow2_64_mod(ctx->hm[k].q);
        hip_check(launch_ntt(ctx->ws0.get<uint32_t>(), count, ctx->dev, scale_fwd(ctx, r64), false, ctx->stream),
                  "ntt forward");
        download_u64(ctx, ctx->ws0.get<uint32_t>(), words, polys);
        return INSPIRE_OK;
    })
}

int inspire_ntt_inverse(inspire_ctx* ctx, uint64_t* polys, uint64_t count) {
    INSPIRE_TRY({
        if (!ctx || (!polys && count)) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        const size_t words = (size_t)count * ctx->K * ctx->N;
        ctx->ws0.reserve(words * 4);
        upload_u64(ctx, polys, words, ctx->ws0.get<uint32_t>());
        uint64_t rinv[kMaxK];  // the reference leaves Montgomery form: times 2^-64
        for (uint32_t k = 0; k < ctx->K; k++) {
            const uint64_t q = ctx->hm[k].q;
            rinv[k] = powmod64(pow2_64_mod(q), q - 2, q);
        }
        hip_check(launch_ntt(ctx->ws0.get<uint32_t>(), count, ctx->dev, scale_inv(ctx, rinv), true, ctx->stream),
                  "ntt inverse");
        download_u64(ctx, ctx->ws0.get<uint32_t>(), words, polys);
        return INSPIRE_OK;
    })
}

int inspire_external_product(inspire_ctx* ctx, uint64_t count, const uint64_t* rlwe, const uint64_t* rgsw,
                             uint64_t* out) {
    INSPIRE_TRY({
        if (!ctx || (count && (!rlwe || !rgsw || !out))) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        external_product_batch(ctx, count, rlwe, rgsw, out);
        return INSPIRE_OK;
    })
}

int inspire_db_load(inspire_ctx* ctx, uint32_t db_id, uint32_t num_shards, const uint32_t* shard_ids,
                    uint32_t polys_per_shard, const uint64_t* polys) {
    INSPIRE_TRY({
        if (!ctx || (num_shards && !shard_ids) || (num_shards && polys_per_shard && !polys))
            throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        owner_only(ctx, "loading");
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        db_load(ctx, db_id, num_shards, shard_ids, polys_per_shard, polys);
        return INSPIRE_OK;
    })
}

int inspire_db_load_entries(inspire_ctx* ctx, uint32_t db_id, const uint8_t* data, uint64_t len,
                            uint32_t entry_size) {
    INSPIRE_TRY({
        if (!ctx || (len && !data)) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        owner_only(ctx, "loading");
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        db_load_entries(ctx, db_id, data, len, entry_size);
        return INSPIRE_OK;
    })
}

int inspire_db_load_shard_files(inspire_ctx* ctx, uint32_t db_id, const char* dir) {
    INSPIRE_TRY({
        if (!ctx || !dir) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        owner_only(ctx, "loading");
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        db_load_shard_files(ctx, db_id, dir);
        return INSPIRE_OK;
    })
}

int inspire_db_unload(inspire_ctx* ctx, uint32_t db_id) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        owner_only(ctx, "unloading");
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        if (!ctx->dbs.erase(db_id)) throw Error{INSPIRE_ERR_UNKNOWN_DATABASE, "unknown database " + std::to_string(db_id)};
        return INSPIRE_OK;
    })
}

static int respond_one(inspire_ctx* ctx, const inspire_query* q, uint32_t variant, inspire_response* r,
                       uint32_t seeded, Route route) {
    INSPIRE_TRY({
        if (!ctx || !q || !r) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (q->seeded != seeded)
            throw Error{INSPIRE_ERR_INVALID_ARGUMENT,
                        seeded ? "this entry point takes a SeededClientQuery (seeded = 1)"
                               : "this entry point takes a ClientQuery (seeded = 0)"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        LaneRead lr(ctx);
        int32_t st = 0;
        respond_batch(ctx, 1, q, &variant, r, &st, route);
        return st;
    })
}

int inspire_respond(inspire_ctx* ctx, const inspire_query* q, inspire_response* r) {
    return respond_one(ctx, q, INSPIRE_VARIANT_NO_PACKING, r, 0, Route::ByVariant);
}

int inspire_respond_seeded(inspire_ctx* ctx, const inspire_query* q, inspire_response* r) {
    return respond_one(ctx, q, INSPIRE_VARIANT_NO_PACKING, r, 1, Route::ByVariant);
}

int inspire_respond_inspiring(inspire_ctx* ctx, const inspire_query* q, inspire_response* r) {
    return respond_one(ctx, q, INSPIRE_VARIANT_ONE_PACKING, r, 0, Route::Inspiring);
}

int inspire_respond_seeded_inspiring(inspire_ctx* ctx, const inspire_query* q, inspire_response* r) {
    return respond_one(ctx, q, INSPIRE_VARIANT_TWO_PACKING, r, 1, Route::Inspiring);
}

int inspire_respond_one_packing(inspire_ctx* ctx, const inspire_query* q, inspire_response* r) {
    return respond_one(ctx, q, INSPIRE_VARIANT_ONE_PACKING, r, 0, Route::Tree);
}

int inspire_respond_seeded_packed(inspire_ctx* ctx, const inspire_query* q, inspire_response* r) {
    return respond_one(ctx, q, INSPIRE_VARIANT_ONE_PACKING, r, 1, Route::Tree);
}

int inspire_ctx_set_galois_keys(inspire_ctx* ctx, const uint64_t* keys, uint32_t count) {
    INSPIRE_TRY({
        if (!ctx || (count && !keys)) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        owner_only(ctx, "loading");
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        set_galois_keys(ctx, keys, count);
        return INSPIRE_OK;
    })
}

int inspire_respond_with_variant(inspire_ctx* ctx, const inspire_query* q, uint32_t variant, inspire_response* r) {
    INSPIRE_TRY({
        if (!ctx || !q || !r) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        LaneRead lr(ctx);
        int32_t st = 0;
        respond_batch(ctx, 1, q, &variant, r, &st, Route::ByVariant);
        return st;
    })
}

int inspire_respond_batch(inspire_ctx* ctx, uint32_t n, const inspire_query* qs, const uint32_t* variants,
                          inspire_response* rs, int32_t* status) {
    INSPIRE_TRY({
        if (!ctx || (n && (!qs || !variants || !rs || !status)))
            throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        LaneRead lr(ctx);
        respond_batch(ctx, n, qs, variants, rs, status, Route::ByVariant);
        return INSPIRE_OK;
    })
}

}  // extern "C"
