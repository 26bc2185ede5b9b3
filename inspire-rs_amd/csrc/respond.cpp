// Batched respond driver (src/pir/respond.rs) and database residency
// (src/pir/setup.rs:105-120, src/pir/encode_db.rs:106-164).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"
#include "kernels.hpp"
#include "modarith.hpp"
#include "pack.hpp"
#include "respond.hpp"

namespace inspire {

void upload_u64(inspire_ctx* ctx, const uint64_t* src, size_t words, uint32_t* dst) {
    if (!words) return;
    ctx->hbuf.resize(words);
    for (size_t i = 0; i < words; i++) ctx->hbuf[i] = (uint32_t)src[i];
    hip_check(hipMemcpyAsync(dst, ctx->hbuf.data(), words * 4, hipMemcpyHostToDevice, ctx->stream), "upload");
    hip_check(hipStreamSynchronize(ctx->stream), "upload sync");
}

void download_u64(inspire_ctx* ctx, const uint32_t* src, size_t words, uint64_t* dst) {
    if (!words) return;
    ctx->hbuf.resize(words);
    hip_check(hipMemcpyAsync(ctx->hbuf.data(), src, words * 4, hipMemcpyDeviceToHost, ctx->stream), "download");
    hip_check(hipStreamSynchronize(ctx->stream), "download sync");
    for (size_t i = 0; i < words; i++) dst[i] = ctx->hbuf[i];
}

// ---------------------------------------------------------------------------
// Databases
// ---------------------------------------------------------------------------

void db_load(inspire_ctx* ctx, uint32_t db_id, uint32_t num_shards, const uint32_t* shard_ids,
             uint32_t polys_per_shard, const uint64_t* polys) {
    Database db;
    db.num_shards = num_shards;
    db.cols = polys_per_shard;
    for (uint32_t s = 0; s < num_shards; s++) db.slot_of.emplace(shard_ids[s], s);  // keeps the first id
    const size_t KN = (size_t)ctx->K * ctx->N;
    const size_t P = (size_t)num_shards * polys_per_shard;
    if (P) {
        db.digits.alloc(P * ctx->ell * KN * 4);
        // Stage in slices so the raw staging buffer stays small.
        const size_t slice = std::max<size_t>(1, (size_t)(64u << 20) / (KN * 4));
        ctx->ws0.reserve(std::min(P, slice) * KN * 4);
        for (size_t p0 = 0; p0 < P; p0 += slice) {
            const size_t np = std::min(slice, P - p0);
            upload_u64(ctx, polys + p0 * KN, np * KN, ctx->ws0.get<uint32_t>());
            uint32_t* dig = db.digits.get<uint32_t>() + p0 * ctx->ell * KN;
            hip_check(launch_db_digits(ctx->ws0.get<uint32_t>(), dig, np, ctx->dev, ctx->stream), "db digits");
            hip_check(launch_ntt(dig, np * ctx->ell, ctx->dev, scale_none(), false, ctx->stream), "db ntt");
        }
        hip_check(hipStreamSynchronize(ctx->stream), "db load");
    }
    ctx->dbs[db_id] = std::move(db);
}

// encode_database (encode_db.rs:106-164) with shard_size = d entries
// (setup.rs:149-153): column c of local entry k is the LE u16 of bytes
// (2c, 2c+1) (extract_column_value, encode_db.rs:169-181).
void db_load_entries(inspire_ctx* ctx, uint32_t db_id, const uint8_t* data, uint64_t len, uint32_t entry_size) {
    if (entry_size == 0 || len == 0) {  // encode_db.rs:112-114: no shards
        db_load(ctx, db_id, 0, nullptr, 0, nullptr);
        return;
    }
    const uint64_t total = len / entry_size, d = ctx->N;
    const uint32_t npolys = (entry_size * 8 + 15) / 16;
    const uint32_t nshards = (uint32_t)((total + d - 1) / d);
    const size_t KN = (size_t)ctx->K * d;
    std::vector<uint32_t> ids(nshards);
    std::vector<uint64_t> polys((size_t)nshards * npolys * KN, 0);
    for (uint32_t s = 0; s < nshards; s++) {
        ids[s] = s;
        const uint64_t off = (uint64_t)s * d, actual = std::min<uint64_t>(d, total - off);
        for (uint32_t c = 0; c < npolys; c++) {
            uint64_t* poly = polys.data() + ((size_t)s * npolys + c) * KN;
            const uint32_t bo = 2 * c;
            for (uint64_t li = 0; li < actual; li++) {
                const uint8_t* e = data + (off + li) * entry_size;
                uint64_t v = 0;
                if (bo + 1 < entry_size) v = (uint64_t)e[bo] | ((uint64_t)e[bo + 1] << 8);
                else if (bo < entry_size) v = e[bo];
                v %= ctx->q;
                for (uint32_t k = 0; k < ctx->K; k++) poly[(size_t)k * d + li] = v % ctx->hm[k].q;
            }
        }
    }
    db_load(ctx, db_id, nshards, ids.data(), npolys, polys.data());
}

// ---------------------------------------------------------------------------
// General external product (config: RLWE with a != 0)
// ---------------------------------------------------------------------------

void external_product_batch(inspire_ctx* ctx, uint64_t count, const uint64_t* rlwe, const uint64_t* rgsw,
                            uint64_t* out) {
    if (!count) return;
    const size_t KN = (size_t)ctx->K * ctx->N, rows = 2 * ctx->ell;
    // workspace: rlwe -> digits [P][2ell][K][N]; rgsw [P][2ell][2][K][N]; out [P][2][K][N]
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(count, 1024));
    ctx->ws0.reserve(chunk * 2 * KN * 4);
    ctx->ws1.reserve(chunk * rows * KN * 4);
    ctx->ws2.reserve(chunk * rows * 2 * KN * 4);
    uint64_t mont[kMaxK];
    for (uint32_t k = 0; k < ctx->K; k++) mont[k] = (1ull << 32) % ctx->hm[k].q;
    for (uint64_t c0 = 0; c0 < count; c0 += chunk) {
        const size_t P = std::min<uint64_t>(chunk, count - c0);
        uint32_t* in = ctx->ws0.get<uint32_t>();
        uint32_t* dig = ctx->ws1.get<uint32_t>();
        uint32_t* G = ctx->ws2.get<uint32_t>();
        upload_u64(ctx, rlwe + c0 * 2 * KN, P * 2 * KN, in);
        upload_u64(ctx, rgsw + c0 * rows * 2 * KN, P * rows * 2 * KN, G);
        hip_check(launch_rlwe_digits(in, dig, P, ctx->dev, ctx->stream), "rlwe digits");
        hip_check(launch_ntt(dig, P * rows, ctx->dev, scale_none(), false, ctx->stream), "digit ntt");
        hip_check(launch_ntt(G, P * rows * 2, ctx->dev, scale_fwd(ctx, mont), false, ctx->stream), "rgsw ntt");
        hip_check(launch_ext_mac(dig, G, in, P, ctx->dev, ctx->stream), "ext mac");
        hip_check(launch_ntt(in, P * 2, ctx->dev, scale_inv(ctx, nullptr), true, ctx->stream), "ext intt");
        download_u64(ctx, in, P * 2 * KN, out + c0 * 2 * KN);
    }
}

// ---------------------------------------------------------------------------
// Respond
// ---------------------------------------------------------------------------

namespace {

using Item = BatchItem;

void fail(int32_t* status, uint32_t i, int code, const std::string& msg) {
    status[i] = code;
    set_last_error(msg);
}

}  // namespace

void respond_batch(inspire_ctx* ctx, uint32_t n, const inspire_query* qs, const uint32_t* variants,
                   inspire_response* rs, int32_t* status, bool direct_inspiring) {
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    // Group runnable queries by database.
    std::map<uint32_t, std::vector<Item>> groups;
    for (uint32_t i = 0; i < n; i++) {
        status[i] = INSPIRE_OK;
        const inspire_query& q = qs[i];
        inspire_response& r = rs[i];
        r.num_column_ciphertexts = 0;
        if (!r.ciphertext) { fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT, "response ciphertext buffer is null"); continue; }
        if (!q.rgsw || (q.seeded && !q.rgsw_seeds)) { fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT, "query rgsw is null"); continue; }
        const uint32_t v = variants[i];
        if (v > INSPIRE_VARIANT_TWO_PACKING) { fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT, "unknown variant"); continue; }
        const bool packed = v != INSPIRE_VARIANT_NO_PACKING;
        if (packed && !direct_inspiring) {  // respond_with_variant (respond.rs:135-149)
            if (q.packing_mode == INSPIRE_PACKING_TREE) {
                fail(status, i, INSPIRE_ERR_UNSUPPORTED, "tree packing (packing_mode=tree) is not on the GPU path");
                continue;
            }
            if (!q.has_packing_keys) {
                fail(status, i, INSPIRE_ERR_PACKING_KEYS_MISSING_TREE,
                     "InspiRING packing keys missing (set packing_mode=tree to use tree packing)");
                continue;
            }
        }
        if (packed && !q.has_packing_keys) {  // respond.rs:296-299
            fail(status, i, INSPIRE_ERR_PACKING_KEYS_MISSING, "InspiRING client packing keys missing from query");
            continue;
        }
        auto it = ctx->dbs.find(q.db_id);
        if (it == ctx->dbs.end()) {
            fail(status, i, INSPIRE_ERR_UNKNOWN_DATABASE, "unknown database " + std::to_string(q.db_id));
            continue;
        }
        const Database& db = it->second;
        auto sit = db.slot_of.find(q.shard_id);
        if (sit == db.slot_of.end()) {  // respond.rs:78-82
            fail(status, i, INSPIRE_ERR_SHARD_NOT_FOUND, "Shard " + std::to_string(q.shard_id) + " not found");
            continue;
        }
        if (db.cols == 0) {  // empty shard (respond.rs:84-90, 307-313): zero ciphertext, [zero]
            std::memset(r.ciphertext, 0, 2 * KN * 8);
            r.num_column_ciphertexts = 1;
            if (r.column_ciphertexts) {
                if (r.column_capacity < 1) { fail(status, i, INSPIRE_ERR_CAPACITY, "column_ciphertexts capacity too small"); continue; }
                std::memset(r.column_ciphertexts, 0, 2 * KN * 8);
            }
            continue;
        }
        if (packed && q.y_all_len == 0 && !q.y_body) {  // respond.rs:365-369
            fail(status, i, INSPIRE_ERR_PACKING_KEYS_EMPTY,
                 "InspiRING packing keys invalid: y_all and y_body are both empty");
            continue;
        }
        if (!packed && r.column_ciphertexts && r.column_capacity < db.cols) {
            fail(status, i, INSPIRE_ERR_CAPACITY, "column_ciphertexts capacity too small");
            continue;
        }
        groups[q.db_id].push_back(Item{i, (int32_t)sit->second, packed});
    }

    uint64_t mont[kMaxK];
    for (uint32_t k = 0; k < ctx->K; k++) mont[k] = (1ull << 32) % ctx->hm[k].q;

    for (auto& kv : groups) {
        const Database& db = ctx->dbs.at(kv.first);
        const uint32_t cols = db.cols;
        for (int pass = 0; pass < 2; pass++) {  // 0: InsPIRe^0, 1: InspiRING
            std::vector<Item> items;
            for (const Item& it : kv.second)
                if (it.packed == (pass == 1)) items.push_back(it);
            if (items.empty()) continue;
            const uint32_t Q = (uint32_t)items.size();
            // ---- stage rows ell..2ell-1: [Q][ell][2][K][N], seeds, slots
            ctx->ws1.reserve((size_t)Q * ell * 2 * KN * 4);
            ctx->ws4.reserve((size_t)Q * (ell * 8 + 1) * 4 + 64);
            uint32_t* rows = ctx->ws1.get<uint32_t>();
            int32_t* d_slot = ctx->ws4.get<int32_t>();
            uint32_t* d_seeds = ctx->ws4.get<uint32_t>() + Q;
            uint32_t* d_err = d_seeds + (size_t)Q * ell * 8;
            std::vector<uint32_t> hrows((size_t)Q * ell * 2 * KN, 0);
            std::vector<uint32_t> hseeds((size_t)Q * ell * 8, 0);
            std::vector<int32_t> hslot(Q);
            bool any_seeded = false;
            for (uint32_t t = 0; t < Q; t++) {
                const inspire_query& q = qs[items[t].idx];
                hslot[t] = items[t].slot;
                for (size_t i = 0; i < ell; i++) {
                    const size_t row = ell + i;  // rows ell..2ell-1 only: the a-digits are 0
                    uint32_t* dst = hrows.data() + ((size_t)t * ell + i) * 2 * KN;
                    if (q.seeded) {
                        any_seeded = true;
                        const uint64_t* b = q.rgsw + row * KN;
                        for (size_t w = 0; w < KN; w++) dst[KN + w] = (uint32_t)b[w];
                        std::memcpy(&hseeds[((size_t)t * ell + i) * 8], q.rgsw_seeds + 32 * row, 32);
                    } else {
                        const uint64_t* ab = q.rgsw + row * 2 * KN;
                        for (size_t w = 0; w < 2 * KN; w++) dst[w] = (uint32_t)ab[w];
                    }
                }
            }
            hip_check(hipMemcpyAsync(rows, hrows.data(), hrows.size() * 4, hipMemcpyHostToDevice, ctx->stream), "rows");
            hip_check(hipMemcpyAsync(d_slot, hslot.data(), Q * 4, hipMemcpyHostToDevice, ctx->stream), "slots");
            hip_check(hipMemsetAsync(d_err, 0, 4, ctx->stream), "err");
            if (any_seeded) {
                hip_check(hipMemcpyAsync(d_seeds, hseeds.data(), hseeds.size() * 4, hipMemcpyHostToDevice, ctx->stream),
                          "seeds");
                // The a-polys of seeded rows come from their seeds: one launch over
                // every row of each run of consecutive seeded queries.
                for (uint32_t t = 0; t < Q;) {
                    if (!qs[items[t].idx].seeded) { t++; continue; }
                    uint32_t e = t;
                    while (e < Q && qs[items[e].idx].seeded) e++;
                    hip_check(launch_expand_seeds(d_seeds + (size_t)t * ell * 8, rows + (size_t)t * ell * 2 * KN,
                                                  (size_t)(e - t) * ell, 2 * KN, d_err, ctx->dev, ctx->stream),
                              "expand seeds");
                    t = e;
                }
            }
            hip_check(launch_ntt(rows, (size_t)Q * ell * 2, ctx->dev, scale_fwd(ctx, mont), false, ctx->stream),
                      "rows ntt");
            if (pass == 0) {
                // ---- InsPIRe^0: [Q][1 + cols][2][K][N]
                const size_t per = (size_t)(1 + cols) * 2 * KN;
                ctx->ws2.reserve(Q * per * 4);
                uint32_t* out = ctx->ws2.get<uint32_t>();
                // polys per query: 2 (combined) + 2 per column; column c at polys 2 + 2c, 3 + 2c
                hip_check(launch_first_dim(db.digits.get<uint32_t>(), rows, d_slot, out, Q, cols, 2 * (1 + cols), 2,
                                           ctx->dev, ctx->stream),
                          "first dim");
                // INTT every poly of the response; the combined slot (polys 0, 1) is
                // overwritten by the fold afterwards.
                hip_check(launch_ntt(out, (size_t)Q * (1 + cols) * 2, ctx->dev, scale_inv(ctx, nullptr), true,
                                     ctx->stream),
                          "cols intt");
                hip_check(launch_sum_cols(out, Q, cols, ctx->dev, ctx->stream), "fold");
                std::vector<uint32_t> hout(Q * per);
                uint32_t herr = 0;
                hip_check(hipMemcpyAsync(hout.data(), out, hout.size() * 4, hipMemcpyDeviceToHost, ctx->stream), "out");
                hip_check(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, ctx->stream), "err");
                hip_check(hipStreamSynchronize(ctx->stream), "respond");
                for (uint32_t t = 0; t < Q; t++) {
                    const uint32_t i = items[t].idx;
                    if (herr) { fail(status, i, INSPIRE_ERR_DEVICE, "seed expansion exhausted its draw budget"); continue; }
                    inspire_response& r = rs[i];
                    const uint32_t* src = hout.data() + t * per;
                    for (size_t w = 0; w < 2 * KN; w++) r.ciphertext[w] = src[w];
                    r.num_column_ciphertexts = cols;
                    if (r.column_ciphertexts)
                        for (size_t w = 0; w < (size_t)cols * 2 * KN; w++) r.column_ciphertexts[w] = src[2 * KN + w];
                }
            } else {
                pack_respond(ctx, db, items.data(), Q, rows, d_slot, d_err, qs, rs, status);
            }
        }
    }
}

}  // namespace inspire
