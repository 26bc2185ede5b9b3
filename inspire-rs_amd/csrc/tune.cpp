// Kernel-variant tuning: measures the variants of the respond path's
// transforms on the device and keeps the faster one per context.
//
//   ntt_fwd_2048             the forward NTT at d = 2048: the LDS radix-8 VALU
//                            kernel or the matrix-core (int8 MFMA) four-step;
//   chain_values_per_thread  the two-modulus packing chain with four values
//                            per thread (radix-4 passes) or eight (radix-8).
//
// Each variant's output is compared with the other's, word for word, on the
// measured inputs; a variant that differs is never chosen.  The choices and
// the measurements persist as a small JSON file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "abi.hpp"
#include "context.hpp"
#include "kernels.hpp"
#include "ntt_mfma.hpp"
#include "pack.hpp"
#include "respond.hpp"

namespace inspire {

namespace {

struct TuneResult {
    double ntt_valu_ms = 0, ntt_mfma_ms = 0;
    bool ntt_mfma_ok = false, ntt_measured = false;
    double chain_v4_ms = 0, chain_v8_ms = 0;
    bool chain_v8_ok = false, chain_measured = false;
    uint32_t ntt_polys = 0, chain_queries = 0;
};

std::string device_name() {
    int dev = 0;
    hipDeviceProp_t p{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "unknown";
    }
    return std::string(p.gcnArchName);
}

std::vector<uint32_t> random_residues(const inspire_ctx* ctx, size_t polys, uint64_t seed) {
    std::mt19937_64 g(seed);
    const size_t N = ctx->N, K = ctx->K;
    std::vector<uint32_t> v(polys * K * N);
    for (size_t p = 0; p < polys; p++)
        for (size_t k = 0; k < K; k++)
            for (size_t i = 0; i < N; i++) v[(p * K + k) * N + i] = (uint32_t)(g() % ctx->hm[k].q);
    return v;
}

// ms per launch of fn over reps launches on the context's stream
template <typename F>
double time_launches(inspire_ctx* ctx, int reps, F&& fn) {
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    fn();  // warm
    hip_check(hipEventRecord(a, ctx->stream), "event");
    for (int r = 0; r < reps; r++) fn();
    hip_check(hipEventRecord(b, ctx->stream), "event");
    hip_check(hipEventSynchronize(b), "event");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, a, b), "event");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / reps;
}

void tune_ntt(inspire_ctx* ctx, TuneResult& R) {
    if (ctx->N != 2048 || !ctx->dev.mfma.tw1) return;
    const uint32_t P = 2048;
    const size_t words = (size_t)P * ctx->K * ctx->N;
    const std::vector<uint32_t> h = random_residues(ctx, P, 0x6e7474);
    DevBuf src(words * 4), d0(words * 4), d1(words * 4);
    hip_check(hipMemcpy(src.get(), h.data(), words * 4, hipMemcpyHostToDevice), "tune upload");
    const size_t KN = (size_t)ctx->K * ctx->N;
    CtxDev C = ctx->dev;
    auto run = [&](uint32_t mfma, DevBuf& dst) {
        C.ntt_mfma = mfma;
        hip_check(launch_ntt_strided(src.get<uint32_t>(), KN, dst.get<uint32_t>(), KN, P, C, scale_none(), false,
                                     ctx->stream),
                  "tune ntt");
    };
    // alternating rounds, the fastest round of each variant kept: a transient
    // slowdown that hits one variant's launches does not decide the choice
    double best[2] = {1e30, 1e30};
    for (int r = 0; r < 5; r++)
        for (uint32_t v = 0; v < 2; v++)
            best[v] = std::min(best[v], time_launches(ctx, 3, [&] { run(v, v ? d1 : d0); }));
    R.ntt_valu_ms = best[0];
    R.ntt_mfma_ms = best[1];
    std::vector<uint32_t> o0(words), o1(words);
    hip_check(hipMemcpy(o0.data(), d0.get(), words * 4, hipMemcpyDeviceToHost), "tune download");
    hip_check(hipMemcpy(o1.data(), d1.get(), words * 4, hipMemcpyDeviceToHost), "tune download");
    R.ntt_mfma_ok = o0 == o1;
    R.ntt_polys = P;
    R.ntt_measured = true;
}

void tune_chain(inspire_ctx* ctx, TuneResult& R) {
    if (!pack_chain_crt_supported(ctx->dev) || ctx->logn != 11) return;
    // a throwaway database of 64 shards of 32-byte entries and 256 seeded
    // InsPIRe^2 queries with y_body keys, all synthetic
    const uint32_t db_id = 0xfffffff0u, Q = 256, shards = 64;
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    std::mt19937_64 g(0x636861696e);
    std::vector<uint8_t> bytes((size_t)shards * ctx->N * 32);
    for (auto& x : bytes) x = (uint8_t)g();
    db_load_entries(ctx, db_id, bytes.data(), bytes.size(), 32);
    const std::vector<uint32_t> b = random_residues(ctx, (size_t)Q * ell, 1), y = random_residues(ctx, (size_t)Q * ell, 2);
    std::vector<uint32_t> seeds((size_t)Q * ell * 8), ids(Q);
    for (auto& s : seeds) s = (uint32_t)g();
    for (auto& i : ids) i = (uint32_t)(g() % shards);
    DevBuf db_(b.size() * 4), dy(y.size() * 4), ds(seeds.size() * 4), o4((size_t)Q * 2 * KN * 4), o8((size_t)Q * 2 * KN * 4);
    hip_check(hipMemcpy(db_.get(), b.data(), b.size() * 4, hipMemcpyHostToDevice), "tune upload");
    hip_check(hipMemcpy(dy.get(), y.data(), y.size() * 4, hipMemcpyHostToDevice), "tune upload");
    hip_check(hipMemcpy(ds.get(), seeds.data(), seeds.size() * 4, hipMemcpyHostToDevice), "tune upload");
    std::vector<int32_t> st(Q);
    inspire_device_batch B{};
    B.count = Q;
    B.db_id = db_id;
    B.variant = INSPIRE_VARIANT_TWO_PACKING;
    B.seeded = 1;
    B.shard_ids = ids.data();
    B.rows = db_.get<uint32_t>();
    B.seeds = ds.get<uint32_t>();
    B.y_body = dy.get<uint32_t>();
    B.status = st.data();
    const uint32_t saved_v = ctx->dev.chain_v;
    const bool prof_on = ctx->prof.on;
    const uint32_t prof_mask = ctx->prof.mask;
    double ms[2] = {0, 0};
    try {
        for (int vi = 0; vi < 2; vi++) {
            ctx->dev.chain_v = vi ? 8 : 4;
            B.out = (vi ? o8 : o4).get<uint32_t>();
            respond_batch_device(ctx, B);  // warm (tables, LDS limits)
            prof_collect(ctx);
            ctx->prof.on = true;
            ctx->prof.mask = 1u << P_CHAIN;
            const double t0 = ctx->prof.total_ms[P_CHAIN];
            const uint64_t n0 = ctx->prof.launches[P_CHAIN];
            for (int r = 0; r < 5; r++) respond_batch_device(ctx, B);
            prof_collect(ctx);
            ms[vi] = (ctx->prof.total_ms[P_CHAIN] - t0) / (double)(ctx->prof.launches[P_CHAIN] - n0);
            ctx->prof.total_ms[P_CHAIN] = t0;
            ctx->prof.launches[P_CHAIN] = n0;
            ctx->prof.on = prof_on;
            ctx->prof.mask = prof_mask;
        }
    } catch (...) {
        ctx->dev.chain_v = saved_v;
        ctx->prof.on = prof_on;
        ctx->prof.mask = prof_mask;
        ctx->dbs.erase(db_id);
        throw;
    }
    ctx->dev.chain_v = saved_v;
    std::vector<uint32_t> h4((size_t)Q * 2 * KN), h8((size_t)Q * 2 * KN);
    hip_check(hipMemcpy(h4.data(), o4.get(), h4.size() * 4, hipMemcpyDeviceToHost), "tune download");
    hip_check(hipMemcpy(h8.data(), o8.get(), h8.size() * 4, hipMemcpyDeviceToHost), "tune download");
    ctx->dbs.erase(db_id);
    R.chain_v4_ms = ms[0];
    R.chain_v8_ms = ms[1];
    R.chain_v8_ok = h4 == h8;
    R.chain_queries = Q;
    R.chain_measured = true;
}

std::string tuning_json(const inspire_ctx* ctx, const TuneResult* R) {
    std::ostringstream o;
    o.precision(6);
    o << "{\"device\": \"" << device_name() << "\", \"ring_dim\": " << ctx->N << ", \"ntt_fwd_2048\": \""
      << (ctx->dev.ntt_mfma ? "mfma" : "valu") << "\", \"chain_values_per_thread\": " << (ctx->dev.chain_v == 8 ? 8 : 4);
    if (R) {
        o << ", \"measured\": {";
        if (R->ntt_measured)
            o << "\"ntt_fwd_2048_valu_ms\": " << R->ntt_valu_ms << ", \"ntt_fwd_2048_mfma_ms\": " << R->ntt_mfma_ms
              << ", \"ntt_polys\": " << R->ntt_polys << ", \"ntt_mfma_bit_exact\": " << (R->ntt_mfma_ok ? "true" : "false")
              << (R->chain_measured ? ", " : "");
        if (R->chain_measured)
            o << "\"chain_v4_ms\": " << R->chain_v4_ms << ", \"chain_v8_ms\": " << R->chain_v8_ms
              << ", \"chain_queries\": " << R->chain_queries
              << ", \"chain_v8_bit_exact\": " << (R->chain_v8_ok ? "true" : "false");
        o << "}";
    }
    o << "}";
    return o.str();
}

// Issue rates of the probes in rates.hip (T operations/s): 8 workgroups of
// 256 threads per CU, each thread 8 chains of 2048 operations, the mean of
// 5 launches after a warm one.
void measure_rates(inspire_ctx* ctx, double* tops, uint32_t n) {
    int dev = 0, cus = 256;
    hipDeviceProp_t p{};
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
    (void)hipGetLastError();
    const uint32_t blocks = 8u * (uint32_t)std::max(cus, 1), iters = 2048;
    DevBuf out((size_t)blocks * 256 * 4);
    for (int kind = 0; kind < 3; kind++) {
        const double ms = time_launches(ctx, 5, [&] {
            hip_check(launch_rate_probe(kind, out.get<uint32_t>(), blocks, iters, ctx->dev.m[0], ctx->stream),
                      "rate probe");
        });
        tops[kind] = (double)rate_ops_per_launch(blocks, iters) / (ms * 1e-3) / 1e12;
    }
    if (n > 3) {  // HBM streaming read: 1 GiB (zeros), TB/s
        const size_t bytes = (size_t)1 << 30;
        DevBuf src(bytes);
        hip_check(hipMemsetAsync(src.get(), 0, bytes, ctx->stream), "read probe fill");
        const double ms = time_launches(ctx, 5, [&] {
            hip_check(launch_read_probe(src.get(), bytes, out.get<uint32_t>(), blocks, ctx->stream), "read probe");
        });
        tops[3] = (double)bytes / (ms * 1e-3) / 1e12;
        hip_check(hipStreamSynchronize(ctx->stream), "read probe");
    }
}

int copy_out(const std::string& s, char* out, uint64_t cap, uint64_t* len) {
    if (len) *len = s.size();
    if (!out) return INSPIRE_OK;
    if (cap < s.size() + 1) throw Error{INSPIRE_ERR_CAPACITY, "tuning report buffer too small"};
    std::memcpy(out, s.c_str(), s.size() + 1);
    return INSPIRE_OK;
}

// the value of "key" in a flat JSON object written by tuning_json
std::string json_field(const std::string& js, const std::string& key) {
    const size_t p = js.find("\"" + key + "\"");
    if (p == std::string::npos) return "";
    size_t c = js.find(':', p);
    if (c == std::string::npos) return "";
    c++;
    while (c < js.size() && (js[c] == ' ' || js[c] == '"')) c++;
    size_t e = c;
    while (e < js.size() && js[e] != '"' && js[e] != ',' && js[e] != '}') e++;
    return js.substr(c, e - c);
}

void apply_choice(inspire_ctx* ctx, const std::string& name, int value) {
    if (name == "ntt_fwd_2048") {
        if (value && !ntt_mfma_supported(ctx->dev))
            throw Error{INSPIRE_ERR_UNSUPPORTED, "the matrix-core NTT needs d = 2048 and moduli below 2^28"};
        ctx->dev.ntt_mfma = value ? 1u : 0u;
    } else if (name == "chain_values_per_thread") {
        if (value != 4 && value != 8) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "chain_values_per_thread is 4 or 8"};
        ctx->dev.chain_v = (uint32_t)value;
    } else {
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "unknown kernel choice '" + name + "'"};
    }
}

}  // namespace

}  // namespace inspire

using namespace inspire;

extern "C" {

int inspire_ctx_autotune(inspire_ctx* ctx, const char* path, char* report, uint64_t capacity, uint64_t* len) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "tuning goes through the lane's owner context"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        graph_cache_clear(ctx->graphs);
        TuneResult R;
        tune_ntt(ctx, R);
        tune_chain(ctx, R);
        if (R.ntt_measured) ctx->dev.ntt_mfma = (R.ntt_mfma_ok && R.ntt_mfma_ms < R.ntt_valu_ms) ? 1u : 0u;
        if (R.chain_measured) ctx->dev.chain_v = (R.chain_v8_ok && R.chain_v8_ms < R.chain_v4_ms) ? 8u : 4u;
        const std::string js = tuning_json(ctx, &R);
        if (path && *path) {
            std::ofstream f(path);
            if (!(f << js << "\n")) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path};
        }
        return copy_out(js, report, capacity, len);
    })
}

int inspire_ctx_load_tuning(inspire_ctx* ctx, const char* path) {
    INSPIRE_TRY({
        if (!ctx || !path) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "tuning goes through the lane's owner context"};
        std::ifstream f(path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string js = ss.str();
        if (!f || js.empty()) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, std::string("cannot read ") + path};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        graph_cache_clear(ctx->graphs);
        const std::string n = json_field(js, "ntt_fwd_2048"), c = json_field(js, "chain_values_per_thread");
        if (!n.empty()) apply_choice(ctx, "ntt_fwd_2048", n == "mfma" ? 1 : 0);
        if (!c.empty()) apply_choice(ctx, "chain_values_per_thread", std::atoi(c.c_str()));
        return INSPIRE_OK;
    })
}

int inspire_ctx_set_kernel_choice(inspire_ctx* ctx, const char* name, int value) {
    INSPIRE_TRY({
        if (!ctx || !name) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "tuning goes through the lane's owner context"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        graph_cache_clear(ctx->graphs);
        apply_choice(ctx, name, value);
        return INSPIRE_OK;
    })
}

int inspire_ctx_measure_rates(inspire_ctx* ctx, double* tops, uint32_t n) {
    INSPIRE_TRY({
        if (!ctx || !tops) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (n < 3) throw Error{INSPIRE_ERR_CAPACITY, "the rate probes report 3 values"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        measure_rates(ctx, tops, n);
        return INSPIRE_OK;
    })
}

int inspire_ctx_get_tuning(const inspire_ctx* ctx, char* out, uint64_t capacity, uint64_t* len) {
    INSPIRE_TRY({
        if (!ctx) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        return copy_out(tuning_json(ctx, nullptr), out, capacity, len);
    })
}

}  // extern "C"
