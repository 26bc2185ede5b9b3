// Serving loop: continuous batching of respond queries on one GPU.
//
// The reference serves queries from concurrent HTTP handlers, each calling
// respond_with_variant / respond_seeded_with_variant on the shared CRS and
// database (src/bin/server.rs:120-236, src/pir/respond.rs:129-183).  Here
// the databases, the CRS-derived packing tables and the clients' packing keys
// stay resident in HBM, and queries are batched as they arrive:
//
//   * every lane owns a HIP stream, device workspaces and pinned staging, and
//     a worker thread;
//   * a free lane takes the queries waiting at that moment (at most
//     max_batch, and only while the device bytes of all queries in flight
//     stay within the HBM budget), stages them with one host-to-device copy,
//     runs them as one batch (queries of any databases and variants
//     together: one launch sequence per variant class and column count),
//     copies the responses back and completes the tickets;
//   * queries that arrive while batches run wait for the next free lane, so
//     several batches overlap on the device.
// A query's response does not depend on the batch it runs in: every kernel
// computes each query from that query's own inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "abi.hpp"
#include "context.hpp"
#include "host_par.hpp"
#include "kernels.hpp"
#include "modarith.hpp"
#include "pack.hpp"
#include "respond.hpp"

namespace inspire {

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct SReq {
    uint64_t ticket = 0;
    uint32_t db_id = 0, shard_id = 0, variant = 0, seeded = 0, client_id = 0;
    std::vector<uint32_t> rows;   // seeded: [ell][K][N] b polys; full: [ell][2][K][N]
    uint32_t seeds[kMaxEll * 8] = {};
    const uint32_t* borrowed = nullptr;  // the caller's rows when borrowed
    const uint8_t* borrowed_seeds = nullptr;
    const uint32_t* rows_ptr() const { return borrowed ? borrowed : rows.data(); }
    const void* seeds_ptr() const { return borrowed_seeds ? (const void*)borrowed_seeds : (const void*)seeds; }
    uint64_t arrival = 0, done_ns = 0;
    int32_t status = INSPIRE_OK;
    std::string err;
    std::unique_ptr<uint32_t[]> out;  // packed: [2][K][N]; InsPIRe^0: [1 + cols][2][K][N]
    uint32_t cols = 0;            // column count of the database (gamma)
    uint64_t dig = 0;             // device address of the shard's digits
    uint64_t yoff = 0;            // word offset of the client's keys in the pool
    size_t footprint = 0;         // device bytes while in flight
    bool done = false;
    bool packed() const { return variant != INSPIRE_VARIANT_NO_PACKING; }
};

struct Lane {
    inspire_ctx* ctx = nullptr;   // lane context: own stream and workspaces
    DevBuf inbox, rows, out;
    PinBuf pin_in, pin_out;
    std::thread th;
};

}  // namespace inspire

struct inspire_server {
    inspire_ctx* ctx = nullptr;
    inspire_server_config cfg{};
    std::vector<std::unique_ptr<inspire::Lane>> lanes;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<inspire::SReq>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<inspire::SReq>> live;
    uint64_t next_ticket = 1;
    uint64_t inflight = 0, budget = 0;
    uint32_t busy = 0;
    bool stop = false;
    inspire_server_stats stats{};
};

namespace inspire {

namespace {

size_t request_footprint(const inspire_ctx* ctx, uint32_t cols, bool packed) {
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    const size_t in = 16 + ell * 32 + ell * 2 * KN * 4;  // staging (at most full rows)
    const size_t rows = ell * 2 * KN * 4;
    // packing workspaces: step-1 sums [cols][K][N] and one b sum per wave [N/64][cols][K] u64
    // (an InsPIRe^0 query holds them too when it shares a packing launch)
    const size_t pack_ws = (size_t)cols * KN * 4 + (size_t)(ctx->N / 64) * cols * ctx->K * 8;
    if (packed) return in + rows + pack_ws + 2 * KN * 4;
    return in + rows + pack_ws + (size_t)(1 + cols) * 2 * KN * 4;
}

void resolve(inspire_server* s, SReq& r);

// One group of a batch: one column count, and one variant class unless
// the packing launches also answer InsPIRe^0 (mixed: per-query roles).
struct Group {
    bool packed = false, mixed = false;
    uint32_t cols = 0;
    std::vector<SReq*> reqs;  // seeded first
    uint32_t S = 0;           // seeded count
    // byte offsets in the staging area / inbox
    size_t o_dig = 0, o_yoff = 0, o_role = 0, o_qout = 0, o_seeds = 0, o_b = 0, o_full = 0, o_err = 0, in_end = 0;
    size_t o_rows = 0;        // in the lane's rows buffer
    size_t o_out = 0;         // in the lane's out buffer
    std::vector<size_t> off;  // each response's word offset from o_out
    size_t out_words = 0;
};

// Words of a request's response block: packed [2][K][N], InsPIRe^0 [1 + cols][2][K][N].
size_t response_words(const inspire_ctx* ctx, const SReq& r) {
    const size_t KN = (size_t)ctx->K * ctx->N;
    return r.packed() ? 2 * KN : (size_t)(1 + r.cols) * 2 * KN;
}

void run_batch(inspire_server* s, Lane& L, std::vector<std::shared_ptr<SReq>>& batch) {
    inspire_ctx* ctx = L.ctx;
    inspire_ctx* own = s->ctx;
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    std::shared_lock<std::shared_mutex> rd(own->rw);

    // key (variant class, cols): 0 InsPIRe^0, 1 packed, 2 both in one launch sequence
    std::map<std::pair<int, uint32_t>, Group> gm;
    for (auto& r : batch) {
        // bind again under this lock: a database may have been reloaded since submit
        resolve(s, *r);
        if (r->status != INSPIRE_OK) continue;
        Group& g = gm[{r->packed() ? 1 : 0, r->cols}];
        g.packed = r->packed();
        g.cols = r->cols;
        g.reqs.push_back(r.get());
    }
    // a column count with both classes: one packing launch sequence with
    // per-query roles, when the fused front end and the two-modulus chain apply
    for (auto it = gm.begin(); it != gm.end();) {
        const uint32_t cols = it->first.second;
        auto pk = gm.find({1, cols});
        if (it->first.first == 0 && pk != gm.end() && pack_mixed_supported(ctx, cols)) {
            Group& m = gm[{2, cols}];
            m.packed = true;
            m.mixed = true;
            m.cols = cols;
            m.reqs = std::move(pk->second.reqs);
            m.reqs.insert(m.reqs.end(), it->second.reqs.begin(), it->second.reqs.end());
            gm.erase(pk);
            it = gm.erase(it);
        } else {
            ++it;
        }
    }
    if (gm.empty()) {  // every query of the batch failed its binding
        const uint64_t t = now_ns();
        for (auto& r : batch) r->done_ns = t;
        return;
    }
    size_t in = 0, rows_words = 0, out_words = 0;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    for (auto& kv : gm) {
        Group& g = kv.second;
        std::stable_partition(g.reqs.begin(), g.reqs.end(), [](const SReq* r) { return r->seeded != 0; });
        const size_t Q = g.reqs.size();
        g.S = (uint32_t)std::count_if(g.reqs.begin(), g.reqs.end(), [](const SReq* r) { return r->seeded != 0; });
        g.off.resize(Q);
        for (size_t t = 0; t < Q; t++) {
            g.off[t] = g.out_words;
            g.out_words += response_words(ctx, *g.reqs[t]);
        }
        g.o_dig = align(in);
        g.o_yoff = align(g.o_dig + Q * 8);
        g.o_role = align(g.o_yoff + Q * 8);
        g.o_qout = align(g.o_role + (g.mixed ? Q * 4 : 0));
        g.o_seeds = align(g.o_qout + (g.mixed ? Q * 8 : 0));
        g.o_b = align(g.o_seeds + (size_t)g.S * ell * 32);
        g.o_full = align(g.o_b + (size_t)g.S * ell * KN * 4);
        g.o_err = align(g.o_full + (Q - g.S) * ell * 2 * KN * 4);
        g.in_end = g.o_err + 4;
        in = g.in_end;
        g.o_rows = rows_words;
        rows_words += Q * ell * 2 * KN;
        g.o_out = out_words;
        out_words += g.out_words;
    }
    L.pin_in.reserve(in);
    L.pin_out.reserve(out_words * 4 + 4 * gm.size() + 16);
    L.inbox.reserve(in);
    L.rows.reserve(rows_words * 4);
    L.out.reserve(out_words * 4);

    // ---- staging (host, on several threads) and one upload
    uint8_t* pin = L.pin_in.get();
    for (auto& kv : gm) {
        Group& g = kv.second;
        const size_t Q = g.reqs.size();
        par_ranges(Q, 8, [&](size_t t0, size_t t1) {
            for (size_t t = t0; t < t1; t++) {
                const SReq* r = g.reqs[t];
                reinterpret_cast<uint64_t*>(pin + g.o_dig)[t] = r->dig;
                reinterpret_cast<uint64_t*>(pin + g.o_yoff)[t] = r->packed() ? r->yoff : 0;
                if (g.mixed) {
                    reinterpret_cast<uint32_t*>(pin + g.o_role)[t] = r->packed() ? 0u : 1u;
                    reinterpret_cast<uint64_t*>(pin + g.o_qout)[t] = g.off[t];
                }
                if (t < g.S) {
                    std::memcpy(pin + g.o_seeds + t * ell * 32, r->seeds_ptr(), ell * 32);
                    std::memcpy(pin + g.o_b + t * ell * KN * 4, r->rows_ptr(), ell * KN * 4);
                } else {
                    std::memcpy(pin + g.o_full + (t - g.S) * ell * 2 * KN * 4, r->rows_ptr(), ell * 2 * KN * 4);
                }
            }
        });
        *reinterpret_cast<uint32_t*>(pin + g.o_err) = 0;
    }
    hipStream_t st = ctx->stream;
    uint8_t* inbox = L.inbox.get<uint8_t>();
    hip_check(hipMemcpyAsync(inbox, pin, in, hipMemcpyHostToDevice, st), "serve upload");

    // ---- per group: rows layout, expansion + NTT, then the respond kernels
    uint32_t* rows = L.rows.get<uint32_t>();
    uint32_t* out = L.out.get<uint32_t>();
    for (auto& kv : gm) {
        Group& g = kv.second;
        const uint32_t Q = (uint32_t)g.reqs.size();
        uint32_t* grows = rows + g.o_rows;
        uint32_t* d_err = reinterpret_cast<uint32_t*>(inbox + g.o_err);
        // packed queries read residue 0 of a only (the InsPIRe^0 ones of a mixed group all of it)
        stage_rows_from(ctx, Q, g.S, grows, reinterpret_cast<const uint32_t*>(inbox + g.o_seeds),
                        reinterpret_cast<const uint32_t*>(inbox + g.o_b),
                        reinterpret_cast<const uint32_t*>(inbox + g.o_full), d_err,
                        g.packed ? pack_rows_a_residues(ctx, g.cols) : 0,
                        g.mixed ? reinterpret_cast<const uint32_t*>(inbox + g.o_role) : nullptr);
        const uint64_t* d_qdig = reinterpret_cast<const uint64_t*>(inbox + g.o_dig);
        PackRoles roles;
        if (g.mixed) {
            roles.role = reinterpret_cast<const uint32_t*>(inbox + g.o_role);
            roles.qout = reinterpret_cast<const uint64_t*>(inbox + g.o_qout);
        }
        if (g.packed)
            pack_run(ctx, g.cols, Q, grows, d_qdig, own->clients.keys.get<uint32_t>(),
                     reinterpret_cast<const uint64_t*>(inbox + g.o_yoff), nullptr, nullptr, out + g.o_out, false,
                     PackWs{}, roles);
        else
            run_respond0(ctx, g.cols, Q, grows, d_qdig, out + g.o_out);
    }
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stats.launch_groups += gm.size();
    }
    uint32_t* pout = L.pin_out.get<uint32_t>();
    hip_check(hipMemcpyAsync(pout, out, out_words * 4, hipMemcpyDeviceToHost, st), "serve download");
    size_t ei = 0;
    for (auto& kv : gm)
        hip_check(hipMemcpyAsync(pout + out_words + ei++, inbox + kv.second.o_err, 4, hipMemcpyDeviceToHost, st),
                  "serve err");
    hip_check(hipStreamSynchronize(st), "serve batch");

    const uint64_t t_done = now_ns();
    for (auto& r : batch) r->done_ns = t_done;
    ei = 0;
    for (auto& kv : gm) {
        Group& g = kv.second;
        const bool err = pout[out_words + ei++] != 0;
        if (err) {
            for (SReq* r : g.reqs) {
                r->status = INSPIRE_ERR_DEVICE;
                r->err = "seed expansion exhausted its draw budget";
            }
            continue;
        }
        // each response out of the lane's pinned buffer (allocated here, copied on several threads)
        for (SReq* r : g.reqs) r->out.reset(new uint32_t[response_words(ctx, *r)]);
        par_ranges(g.reqs.size(), 8, [&](size_t t0, size_t t1) {
            for (size_t t = t0; t < t1; t++)
                std::memcpy(g.reqs[t]->out.get(), pout + g.o_out + g.off[t], response_words(ctx, *g.reqs[t]) * 4);
        });
    }
}

void worker(inspire_server* s, Lane* L) {
    for (;;) {
        std::vector<std::shared_ptr<SReq>> batch;
        size_t fp = 0;
        {
            std::unique_lock<std::mutex> lk(s->m);
            s->cv_work.wait(lk, [&] {
                return (s->stop && s->pending.empty()) ||
                       (!s->pending.empty() && s->inflight + s->pending.front()->footprint <= s->budget);
            });
            if (s->pending.empty()) return;  // stopping
            if (s->cfg.max_wait_us && s->busy > 0 && s->pending.size() < s->cfg.max_batch) {
                s->cv_work.wait_for(lk, std::chrono::microseconds(s->cfg.max_wait_us),
                                    [&] { return s->stop || s->pending.size() >= s->cfg.max_batch; });
                if (s->pending.empty() || s->inflight + s->pending.front()->footprint > s->budget) continue;
            }
            while (!s->pending.empty() && batch.size() < s->cfg.max_batch &&
                   s->inflight + fp + s->pending.front()->footprint <= s->budget) {
                fp += s->pending.front()->footprint;
                batch.push_back(std::move(s->pending.front()));
                s->pending.pop_front();
            }
            s->inflight += fp;
            s->busy++;
            s->stats.batches++;
            s->stats.max_batch_seen = std::max<uint64_t>(s->stats.max_batch_seen, batch.size());
            s->stats.peak_inflight_bytes = std::max<uint64_t>(s->stats.peak_inflight_bytes, s->inflight);
            s->stats.peak_lanes_busy = std::max<uint64_t>(s->stats.peak_lanes_busy, s->busy);
        }
        try {
            run_batch(s, *L, batch);
        } catch (const Error& e) {
            const uint64_t t = now_ns();
            for (auto& r : batch) {
                r->status = e.code;
                r->err = e.msg;
                r->done_ns = t;
            }
        } catch (const std::exception& e) {
            const uint64_t t = now_ns();
            for (auto& r : batch) {
                r->status = INSPIRE_ERR_DEVICE;
                r->err = e.what();
                r->done_ns = t;
            }
        }
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->inflight -= fp;
            s->busy--;
            for (auto& r : batch) r->done = true;
            s->stats.requests += batch.size();
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

// Resolves a request against the resident state (shared lock held): its
// shard's digits, column count, client keys and footprint.  Failures become
// the request's status (the reference's messages).
void resolve(inspire_server* s, SReq& r) {
    if (r.status != INSPIRE_OK) return;
    inspire_ctx* ctx = s->ctx;
    if (r.variant > INSPIRE_VARIANT_TWO_PACKING) {
        r.status = INSPIRE_ERR_INVALID_ARGUMENT;
        r.err = "unknown variant";
        return;
    }
    auto it = ctx->dbs.find(r.db_id);
    if (it == ctx->dbs.end()) {
        r.status = INSPIRE_ERR_UNKNOWN_DATABASE;
        r.err = "unknown database " + std::to_string(r.db_id);
        return;
    }
    const Database& db = it->second;
    if (r.packed()) {
        auto c = ctx->clients.slot_of.find(r.client_id);
        if (c == ctx->clients.slot_of.end()) {  // respond.rs:296-299
            r.status = INSPIRE_ERR_PACKING_KEYS_MISSING;
            r.err = "InspiRING client packing keys missing from query";
            return;
        }
        r.yoff = (uint64_t)c->second * ctx->ell * ctx->K * ctx->N;
    }
    const int64_t slot = db.slot(r.shard_id);
    if (slot < 0) {  // respond.rs:78-82
        r.status = INSPIRE_ERR_SHARD_NOT_FOUND;
        r.err = "Shard " + std::to_string(r.shard_id) + " not found";
        return;
    }
    if (db.cols == 0) {
        r.status = INSPIRE_ERR_UNSUPPORTED;
        r.err = "the serving loop needs a database with columns";
        return;
    }
    if (r.packed()) {
        // the lanes rotate the registered y_body by the NTT-slot permutation of
        // tau_{g^i}, which is generate_rotations only for odd g^i
        std::lock_guard<std::mutex> lk(ctx->packs_mu);
        auto t = ctx->packs.find(db.cols);
        if (t == ctx->packs.end()) {
            r.status = INSPIRE_ERR_UNSUPPORTED;
            r.err = "no packing tables for this database (loaded after the server started)";
            return;
        }
        if (!t->second->odd_gens) {
            r.status = INSPIRE_ERR_UNSUPPORTED;
            r.err = kEvenGeneratorMsg;
            return;
        }
    }
    r.cols = db.cols;
    r.dig = shard_digits(db, (uint32_t)slot);
    r.footprint = request_footprint(ctx, db.cols, r.packed());
    if (r.footprint > s->budget) {  // could never be admitted
        r.status = INSPIRE_ERR_CAPACITY;
        r.err = "query footprint exceeds the server's HBM budget";
    }
}

}  // namespace

}  // namespace inspire

using namespace inspire;

extern "C" {

int inspire_register_client_keys(inspire_ctx* ctx, uint32_t client_id, const uint64_t* y_body) {
    INSPIRE_TRY({
        if (!ctx || !y_body) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "registering keys goes through the lane's owner context"};
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::unique_lock<std::shared_mutex> w(ctx->rw);
        const size_t KN = (size_t)ctx->K * ctx->N, per = ctx->ell * KN;
        ClientPool& P = ctx->clients;
        uint32_t slot;
        auto it = P.slot_of.find(client_id);
        if (it != P.slot_of.end()) {
            slot = it->second;
        } else {
            if (P.used == P.capacity) {  // grow, keeping the registered keys
                const uint32_t cap = std::max<uint32_t>(64, 2 * P.capacity);
                DevBuf grown(cap * per * 4);
                if (P.used)
                    hip_check(hipMemcpyAsync(grown.get(), P.keys.get(), (size_t)P.used * per * 4,
                                             hipMemcpyDeviceToDevice, ctx->stream),
                              "client keys grow");
                hip_check(hipStreamSynchronize(ctx->stream), "client keys grow");
                P.keys = std::move(grown);
                P.capacity = cap;
            }
            slot = P.used++;
        }
        uint32_t* dst = P.keys.get<uint32_t>() + (size_t)slot * per;
        upload_u64(ctx, y_body, per, dst);
        uint64_t mont[kMaxK];
        for (uint32_t k = 0; k < ctx->K; k++) mont[k] = (1ull << 32) % ctx->hm[k].q;
        hip_check(launch_ntt(dst, ctx->ell, ctx->dev, scale_fwd(ctx, mont), false, ctx->stream), "client keys ntt");
        hip_check(hipStreamSynchronize(ctx->stream), "client keys");
        P.slot_of[client_id] = slot;
        return INSPIRE_OK;
    })
}

int inspire_server_create(inspire_ctx* ctx, const inspire_server_config* cfg, inspire_server** out) {
    INSPIRE_TRY({
        if (!ctx || !cfg || !out) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        *out = nullptr;
        if (ctx->owner) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "a serving lane cannot host a server"};
        if (cfg->lanes < 1 || cfg->lanes > 8) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "lanes must be in [1, 8]"};
        if (cfg->max_batch < 1 || cfg->max_batch > 1024)
            throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "max_batch must be in [1, 1024]"};
        std::unique_ptr<inspire_server> s(new inspire_server());
        s->ctx = ctx;
        s->cfg = *cfg;
        size_t max_fp = request_footprint(ctx, 1, false);
        {
            std::lock_guard<std::recursive_mutex> g(ctx->mu);
            std::unique_lock<std::shared_mutex> w(ctx->rw);
            for (auto& kv : ctx->dbs) {
                const uint32_t cols = kv.second.cols;
                if (!cols) continue;
                max_fp = std::max({max_fp, request_footprint(ctx, cols, false), request_footprint(ctx, cols, true)});
                try {
                    pack_tables(ctx, cols);  // the lanes only read tables
                } catch (const Error&) {
                    // a column count InspiRING cannot pack: packed queries to it fail at submit
                }
            }
        }
        s->budget = cfg->hbm_budget ? std::max<uint64_t>(cfg->hbm_budget, max_fp)
                                    : (uint64_t)cfg->lanes * cfg->max_batch * max_fp;
        try {
            for (uint32_t i = 0; i < cfg->lanes; i++) {
                auto L = std::make_unique<Lane>();
                L->ctx = make_lane_ctx(ctx);
                s->lanes.push_back(std::move(L));
            }
        } catch (...) {
            for (auto& L : s->lanes) destroy_lane_ctx(L->ctx);
            throw;
        }
        for (auto& L : s->lanes) L->th = std::thread(worker, s.get(), L.get());
        *out = s.release();
        return INSPIRE_OK;
    })
}

int inspire_server_submit(inspire_server* s, uint32_t n, const inspire_request* reqs, uint64_t* tickets) {
    INSPIRE_TRY({
        if (!s || (n && (!reqs || !tickets))) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        inspire_ctx* ctx = s->ctx;
        const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
        const uint64_t t_now = now_ns();
        std::vector<std::shared_ptr<SReq>> made(n);
        {
            std::shared_lock<std::shared_mutex> rd(ctx->rw);
            for (uint32_t i = 0; i < n; i++) {
                const inspire_request& q = reqs[i];
                auto r = std::make_shared<SReq>();
                r->db_id = q.db_id;
                r->shard_id = q.shard_id;
                r->variant = q.variant;
                r->seeded = q.seeded ? 1 : 0;
                r->client_id = q.client_id;
                r->arrival = q.arrival_ns ? q.arrival_ns : t_now;
                if (!q.rows || (q.seeded && !q.seeds)) {
                    r->status = INSPIRE_ERR_INVALID_ARGUMENT;
                    r->err = "request rows or seeds are null";
                } else {
                    resolve(s, *r);
                }
                if (r->status == INSPIRE_OK) {
                    if (q.flags & INSPIRE_REQUEST_BORROW) {
                        r->borrowed = q.rows;
                        r->borrowed_seeds = r->seeded ? q.seeds : nullptr;
                    } else {
                        const size_t words = (r->seeded ? 1 : 2) * ell * KN;
                        r->rows.assign(q.rows, q.rows + words);
                        if (r->seeded) std::memcpy(r->seeds, q.seeds, ell * 32);
                    }
                }
                made[i] = std::move(r);
            }
        }
        {
            std::lock_guard<std::mutex> lk(s->m);
            if (s->stop) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "the server is stopping"};
            for (uint32_t i = 0; i < n; i++) {
                auto& r = made[i];
                r->ticket = s->next_ticket++;
                tickets[i] = r->ticket;
                s->live[r->ticket] = r;
                if (r->status == INSPIRE_OK) {
                    s->pending.push_back(r);
                } else {
                    r->done = true;
                    r->done_ns = t_now;
                }
            }
        }
        s->cv_work.notify_all();
        s->cv_done.notify_all();
        return INSPIRE_OK;
    })
}

}  // extern "C"

namespace inspire {
namespace {

// Waits for one ticket and releases it (inspire_server_wait).
int wait_ticket(inspire_server* s, uint64_t ticket, int32_t timeout_ms, inspire_response* r, int32_t* status,
                uint64_t* done_ns) {
    std::shared_ptr<SReq> q;
    {
        std::unique_lock<std::mutex> lk(s->m);
        auto it = s->live.find(ticket);
        if (it == s->live.end()) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "unknown ticket"};
        q = it->second;
        auto ready = [&] { return q->done; };
        if (timeout_ms < 0) s->cv_done.wait(lk, ready);
        else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
            set_last_error("ticket not finished yet");
            return INSPIRE_ERR_PENDING;
        }
        s->live.erase(ticket);  // the map may have changed while waiting
    }
    if (status) *status = q->status;
    if (done_ns) *done_ns = q->done_ns;
    if (q->status != INSPIRE_OK) set_last_error(q->err);
    if (r && q->status == INSPIRE_OK) {
        const size_t KN = (size_t)s->ctx->K * s->ctx->N;
        if (!r->ciphertext) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "response ciphertext buffer is null"};
        r->ciphertext_b_dim = s->ctx->N;
        for (size_t w = 0; w < 2 * KN; w++) r->ciphertext[w] = q->out[w];
        if (q->packed()) {
            r->num_column_ciphertexts = 0;  // respond.rs:382-385
        } else {
            r->num_column_ciphertexts = q->cols;
            if (r->column_ciphertexts) {
                if (r->column_capacity < q->cols)
                    throw Error{INSPIRE_ERR_CAPACITY, "column_ciphertexts capacity too small"};
                for (size_t w = 0; w < (size_t)q->cols * 2 * KN; w++) r->column_ciphertexts[w] = q->out[2 * KN + w];
            }
        }
    }
    return INSPIRE_OK;
}

}  // namespace
}  // namespace inspire

extern "C" {

int inspire_server_wait(inspire_server* s, uint64_t ticket, int32_t timeout_ms, inspire_response* r, int32_t* status,
                        uint64_t* done_ns) {
    INSPIRE_TRY({
        if (!s) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        return wait_ticket(s, ticket, timeout_ms, r, status, done_ns);
    })
}

int inspire_server_wait_many(inspire_server* s, uint32_t n, const uint64_t* tickets, int32_t timeout_ms,
                             inspire_response* rs, int32_t* status, uint64_t* done_ns) {
    INSPIRE_TRY({
        if (!s || (n && (!tickets || !status))) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        int rc = INSPIRE_OK;
        for (uint32_t i = 0; i < n; i++) {
            const int c = wait_ticket(s, tickets[i], timeout_ms, rs ? rs + i : nullptr, status + i,
                                      done_ns ? done_ns + i : nullptr);
            if (c == INSPIRE_ERR_PENDING) {
                status[i] = INSPIRE_ERR_PENDING;
                rc = INSPIRE_ERR_PENDING;
            }
        }
        return rc;
    })
}

int inspire_server_get_stats(inspire_server* s, inspire_server_stats* out) {
    INSPIRE_TRY({
        if (!s || !out) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument"};
        std::lock_guard<std::mutex> lk(s->m);
        *out = s->stats;
        return INSPIRE_OK;
    })
}

void inspire_server_destroy(inspire_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& L : s->lanes)
        if (L->th.joinable()) L->th.join();
    for (auto& L : s->lanes) {
        L->inbox.release();
        L->rows.release();
        L->out.release();
        destroy_lane_ctx(L->ctx);
    }
    delete s;
}

}  // extern "C"
