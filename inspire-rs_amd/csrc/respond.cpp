// Batched respond driver (src/pir/respond.rs) and database residency
// (src/pir/setup.rs:105-120, src/pir/encode_db.rs:106-164).
//
// A respond batch is split by database and by variant class (InsPIRe^0 or
// InspiRING); each group runs the same device stages whether its inputs came
// from host buffers (the reference-shaped API) or are already resident in
// HBM (inspire_respond_batch_device):
//   rows   RGSW rows ell..2ell-1 -> (seed expansion) -> NTT, Montgomery form
//   ^0     first dimension -> INTT of every column -> column fold
//   ^1/^2  first dimension -> pack (b_poly, step 1, chain): pack.cpp
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "context.hpp"
#include "host_par.hpp"
#include "kernels.hpp"
#include "modarith.hpp"
#include "pack.hpp"
#include "respond.hpp"
#include "tree.hpp"

namespace inspire {

const char* const kProfNames[P_COUNT] = {"rows_ntt", "expand_seeds", "first_dim", "cols_intt", "fold",
                                         "pack_b0",  "pack_step1",   "y_ntt",     "pack_chain", "db_load",
                                         "pack_front", "tree_leaves", "tree_combine", "extract"};

ProfScope::ProfScope(inspire_ctx* ctx, int id) : ctx_(ctx), id_(id) {
    if (!ctx_->prof.on || ctx_->prof.suspended || !((ctx_->prof.mask >> id) & 1u)) return;
    if (ctx_->graphs.segmenting) {
        seg_ = true;
        graph_segment(ctx_, -1);
        return;
    }
    auto& P = ctx_->prof;
    if (P.pool.empty()) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "event");
        P.pool.push_back(e);
    }
    a_ = P.pool.back();
    P.pool.pop_back();
    hip_check(hipEventRecord(a_, ctx_->stream), "event record");
}

ProfScope::~ProfScope() {
    if (seg_) graph_segment(ctx_, id_);
    if (!a_) return;
    auto& P = ctx_->prof;
    hipEvent_t b;
    if (P.pool.empty()) {
        if (hipEventCreate(&b) != hipSuccess) { (void)hipGetLastError(); return; }
    } else {
        b = P.pool.back();
        P.pool.pop_back();
    }
    (void)hipEventRecord(b, ctx_->stream);
    P.spans.push_back({id_, a_, b});
}

static bool prof_wants(const inspire_ctx* ctx, int id) {
    return ctx->prof.on && !ctx->prof.suspended && ((ctx->prof.mask >> id) & 1u);
}

// The launch with two events from the pool, its span kept for prof_collect.
static void timed_launch(inspire_ctx* ctx, int id, const char* what, const LaunchFn& launch) {
    auto& P = ctx->prof;
    hipEvent_t e[2];
    for (int j = 0; j < 2; j++) {
        if (P.pool.empty()) {
            hip_check(hipEventCreate(&e[j]), "event");
        } else {
            e[j] = P.pool.back();
            P.pool.pop_back();
        }
    }
    const hipError_t err = launch(LaunchEv{e[0], e[1]});
    if (err != hipSuccess) {
        P.pool.push_back(e[0]);
        P.pool.push_back(e[1]);
        hip_check(err, what);
    }
    P.spans.push_back({id, e[0], e[1]});
}

bool chain_order_on(const inspire_ctx* ctx) { return root(ctx)->chain_window_us.load() > 0 && ctx->chain_done; }

// `launch`, after the owner's previous chain launch when that one was
// submitted from another context less than the window before (a wait on its
// event in this stream, no host synchronisation); this context's event is
// then recorded as the last.  Lanes that submit together would otherwise run
// their chains together, each beside the others' (every workgroup of a chain
// fills a compute unit), and keep doing so batch after batch, their front
// ends idle meanwhile; chains submitted further apart overlap freely.
static LaunchFn in_chain_order(inspire_ctx* ctx, LaunchFn launch) {
    return [ctx, launch = std::move(launch)](LaunchEv ev) -> hipError_t {
        inspire_ctx* R = root(ctx);
        std::lock_guard<std::mutex> lk(R->chain_mu);
        const auto now = std::chrono::steady_clock::now();
        const auto window = std::chrono::microseconds(R->chain_window_us.load());
        // The order only schedules: when this runtime refuses the wait (it does
        // while the stream the event was recorded on is capturing a graph), or
        // the record, the chain runs unordered and the result is the same.
        if (R->last_chain && R->last_chain != ctx->chain_done && now - R->last_submit < window &&
            hipStreamWaitEvent(ctx->stream, R->last_chain, 0) != hipSuccess)
            (void)hipGetLastError();
        const hipError_t e = launch(ev);
        if (e != hipSuccess) return e;
        if (hipEventRecord(ctx->chain_done, ctx->stream) == hipSuccess) {
            R->last_chain = ctx->chain_done;
            R->last_submit = now;
        } else {
            (void)hipGetLastError();
            R->last_chain = nullptr;
        }
        return hipSuccess;
    };
}

void prof_launch(inspire_ctx* ctx, int id, const char* what, LaunchFn launch, bool ordered) {
    ordered = ordered && chain_order_on(ctx);
    if (ordered) launch = in_chain_order(ctx, std::move(launch));
    const bool timed = prof_wants(ctx, id);
    GraphCache& G = ctx->graphs;
    if (G.segmenting && (timed || ordered)) {
        graph_segment(ctx, -1);  // the graph segment before it
        if (G.seg_error == hipSuccess) {
            GraphSeg sg;
            sg.prof = id;
            sg.eager = std::move(launch);
            sg.what = what;
            G.pending.push_back(std::move(sg));
        }
        return;
    }
    if (timed) timed_launch(ctx, id, what, launch);
    else hip_check(launch(LaunchEv{}), what);
}

void prof_collect(inspire_ctx* ctx) {
    auto& P = ctx->prof;
    if (P.spans.empty()) return;
    hip_check(hipStreamSynchronize(ctx->stream), "profile sync");
    for (auto& s : P.spans) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) {
            P.total_ms[s.id] += ms;
            P.launches[s.id] += 1;
        } else {
            (void)hipGetLastError();
        }
        P.pool.push_back(s.a);
        P.pool.push_back(s.b);
    }
    P.spans.clear();
}

static void destroy_segs(std::vector<GraphSeg>& segs) {
    for (auto& sg : segs) {
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
        if (sg.graph) (void)hipGraphDestroy(sg.graph);
    }
    segs.clear();
}

void graph_cache_clear(GraphCache& G) {
    for (auto& e : G.entries) destroy_segs(e.segs);
    G.entries.clear();
    G.last_eager.clear();
}

// Thread-local: other threads' calls are not restricted by this capture.
// (This runtime still refuses null-stream operations of other threads while
// any capture is open, with hipErrorStreamCaptureImplicit; callers that copy
// concurrently use streams of their own.  In relaxed mode such a copy crashed
// the runtime instead.)
constexpr hipStreamCaptureMode kCaptureMode = hipStreamCaptureModeThreadLocal;

void graph_segment(inspire_ctx* ctx, int prof) noexcept {
    GraphCache& G = ctx->graphs;
    if (G.seg_error != hipSuccess) return;  // the capture is abandoned after the call
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    size_t nodes = 0;
    if (e == hipSuccess) e = hipGraphGetNodes(g, nullptr, &nodes);
    if (e == hipSuccess && nodes) G.pending.push_back(GraphSeg{g, nullptr, prof});
    else if (g) (void)hipGraphDestroy(g);
    if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, kCaptureMode);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        G.seg_error = e;
    }
}

namespace {

constexpr size_t kGraphCacheEntries = 8;

// Launches a captured sequence: graph segments, a profiled one between two
// events recorded on the stream, and the profiled launches kept out of the
// graphs, each with its own events.
void launch_segs(inspire_ctx* ctx, const GraphEntry& e) {
    for (const auto& sg : e.segs) {
        if (sg.eager) {  // a profiled or ordered launch
            if (prof_wants(ctx, sg.prof)) timed_launch(ctx, sg.prof, sg.what, sg.eager);
            else hip_check(sg.eager(LaunchEv{}), sg.what);
        } else if (sg.prof >= 0) {
            ProfScope ps(ctx, sg.prof);
            hip_check(hipGraphLaunch(sg.exec, ctx->stream), "graph launch");
        } else {
            hip_check(hipGraphLaunch(sg.exec, ctx->stream), "graph launch");
        }
    }
}

// Runs `enqueue` (every stream operation of one device batch) eagerly, or
// replays the graph captured for `key`.  A key seen on one of the last
// eager calls is captured now, so one-off shapes never pay for a capture.  When
// capture or instantiation is refused the call runs eagerly.  With profiling
// on, the profiled kernels split the capture into segments (the key carries
// the profiled set).
template <typename F>
void run_with_graphs(inspire_ctx* ctx, const std::vector<uint64_t>& key, F&& enqueue) {
    GraphCache& G = ctx->graphs;
    for (size_t i = 0; i < G.entries.size(); i++) {
        if (G.entries[i].key != key) continue;
        GraphEntry e = std::move(G.entries[i]);  // most recently used goes last
        G.entries.erase(G.entries.begin() + (ptrdiff_t)i);
        G.entries.push_back(std::move(e));
        G.replays++;
        launch_segs(ctx, G.entries.back());
        return;
    }
    auto seen = std::find(G.last_eager.begin(), G.last_eager.end(), key);
    if (seen == G.last_eager.end()) {
        enqueue();
        G.last_eager.push_back(key);
        if (G.last_eager.size() > kGraphCacheEntries) G.last_eager.erase(G.last_eager.begin());
        return;
    }
    G.last_eager.erase(seen);
    if (hipStreamBeginCapture(ctx->stream, kCaptureMode) != hipSuccess) {
        (void)hipGetLastError();
        G.enabled = false;  // this runtime cannot capture: stay eager
        enqueue();
        return;
    }
    destroy_segs(G.pending);
    G.seg_error = hipSuccess;
    G.segmenting = (ctx->prof.on && !ctx->prof.suspended) || chain_order_on(ctx);
    try {
        enqueue();
    } catch (...) {
        // the capture failed (invalidated, or an operation refused while it
        // was open): drop it and run this call eagerly, which reports any
        // error of its own
        G.segmenting = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        (void)hipGetLastError();
        if (g) (void)hipGraphDestroy(g);
        destroy_segs(G.pending);
        enqueue();
        return;
    }
    G.segmenting = false;
    graph_segment(ctx, -1);  // closes the last segment (and begins one more, ended here)
    bool ok = G.seg_error == hipSuccess;
    {
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(ctx->stream, &g) != hipSuccess) (void)hipGetLastError();
        if (g) (void)hipGraphDestroy(g);
    }
    GraphEntry e;
    e.key = key;
    e.segs = std::move(G.pending);
    G.pending.clear();
    for (auto& sg : e.segs) {
        if (!ok) break;
        if (!sg.graph) continue;  // a launch kept out of the graphs
        if (hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            sg.exec = nullptr;
            ok = false;
        }
    }
    if (!ok) {  // run this call eagerly
        destroy_segs(e.segs);
        enqueue();
        return;
    }
    if (G.entries.size() >= kGraphCacheEntries) {  // evict the least recently used
        destroy_segs(G.entries.front().segs);
        G.entries.erase(G.entries.begin());
    }
    G.entries.push_back(std::move(e));
    G.captures++;
    launch_segs(ctx, G.entries.back());
}

}  // namespace

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

static uint64_t mont_of(const inspire_ctx* ctx, uint32_t k) { return (1ull << 32) % ctx->hm[k].q; }

// ---------------------------------------------------------------------------
// Databases
// ---------------------------------------------------------------------------

// Preprocesses `np` shard-column polys of residues (device, [np][K][N]) into
// NTT-domain gadget digits of Delta*h at `dig` ([np][ell][K][N]).
static void db_digits(inspire_ctx* ctx, const uint32_t* h, size_t np, uint32_t* dig) {
    ProfScope ps(ctx, P_DB);
    hip_check(launch_db_digits(h, dig, np, ctx->dev, ctx->stream), "db digits");
    hip_check(launch_ntt(dig, np * ctx->ell, ctx->dev, scale_none(), false, ctx->stream), "db ntt");
}

uint64_t shard_digits(const Database& db, uint32_t slot) {
    return reinterpret_cast<uint64_t>(db.digits.get<uint32_t>() + (size_t)slot * db.digit_words_per_shard);
}

static void db_begin(const inspire_ctx* ctx, Database& db, uint32_t num_shards, const uint32_t* shard_ids,
                     uint32_t cols) {
    db.num_shards = num_shards;
    db.cols = cols;
    db.digit_words_per_shard = (size_t)cols * ctx->ell * ctx->K * ctx->N;
    db.dense_ids = true;
    for (uint32_t s = 0; s < num_shards; s++) {
        const uint32_t id = shard_ids ? shard_ids[s] : s;
        db.dense_ids = db.dense_ids && id == s;
    }
    if (!db.dense_ids)
        for (uint32_t s = 0; s < num_shards; s++) db.slot_of.emplace(shard_ids[s], s);  // first id wins
}

void db_load(inspire_ctx* ctx, uint32_t db_id, uint32_t num_shards, const uint32_t* shard_ids,
             uint32_t polys_per_shard, const uint64_t* polys) {
    Database db;
    db_begin(ctx, db, num_shards, shard_ids, polys_per_shard);
    const size_t KN = (size_t)ctx->K * ctx->N;
    const size_t P = (size_t)num_shards * polys_per_shard;
    ctx->dbs.erase(db_id);  // release the old copy before allocating the new one
    if (P) {
        db.digits.alloc(P * ctx->ell * KN * 4);
        const size_t slice = std::max<size_t>(1, (size_t)(64u << 20) / (KN * 4));
        ctx->ws0.reserve(std::min(P, slice) * KN * 4);
        for (size_t p0 = 0; p0 < P; p0 += slice) {
            const size_t np = std::min(slice, P - p0);
            upload_u64(ctx, polys + p0 * KN, np * KN, ctx->ws0.get<uint32_t>());
            db_digits(ctx, ctx->ws0.get<uint32_t>(), np, db.digits.get<uint32_t>() + p0 * ctx->ell * KN);
        }
        hip_check(hipStreamSynchronize(ctx->stream), "db load");
    }
    prof_collect(ctx);
    ctx->dbs[db_id] = std::move(db);
}

// load_shard_binary (src/pir/mmap.rs:55-85) of every shard-*.bin file in a
// directory (the layout save_shards_binary writes, mmap.rs:19-52): LE u32 id,
// u32 polynomial count, then (count > 0) u32 ring_dim, u32 crt_count and u64
// moduli, then the polynomials' residue-major u64 coefficients.  The shards
// are ingested into device memory once instead of being read per query
// (respond_mmap*, respond.rs:492-640).
void db_load_shard_files(inspire_ctx* ctx, uint32_t db_id, const char* dir) {
    namespace fs = std::filesystem;
    struct File {
        std::string path;
        uint32_t id = 0, polys = 0;
        size_t data_off = 0;
    };
    std::vector<File> files;
    std::error_code ec;
    fs::directory_iterator it(dir, ec);
    if (ec) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, std::string("cannot read shard directory ") + dir};
    for (const auto& e : it) {
        const std::string name = e.path().filename().string();
        if (name.rfind("shard-", 0) == 0 && name.size() > 4 && name.compare(name.size() - 4, 4, ".bin") == 0)
            files.push_back(File{e.path().string()});
    }
    const size_t KN = (size_t)ctx->K * ctx->N;
    auto bad = [](const std::string& f, const std::string& why) {
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "shard file " + f + ": " + why};
    };
    uint32_t cols = 0;
    bool first = true;
    for (File& f : files) {
        std::ifstream in(f.path, std::ios::binary);
        uint32_t h[4] = {0, 0, 0, 0};
        if (!in.read(reinterpret_cast<char*>(h), 8)) bad(f.path, "truncated header");
        f.id = h[0];
        f.polys = h[1];
        if (!in.read(reinterpret_cast<char*>(h + 2), 8)) bad(f.path, "truncated header");
        f.data_off = 16;
        if (f.polys) {
            if (h[2] != ctx->N) bad(f.path, "ring dimension " + std::to_string(h[2]) + " differs from the context's");
            if (h[3] != ctx->K) bad(f.path, "CRT modulus count " + std::to_string(h[3]) + " differs from the context's");
            for (uint32_t k = 0; k < ctx->K; k++) {
                uint64_t m = 0;
                if (!in.read(reinterpret_cast<char*>(&m), 8)) bad(f.path, "truncated header");
                if (m != ctx->hm[k].q) bad(f.path, "CRT moduli differ from the context's");
            }
            f.data_off += 8ull * ctx->K;
            const uint64_t want = f.data_off + (uint64_t)f.polys * KN * 8;
            if (fs::file_size(f.path, ec) < want || ec) bad(f.path, "truncated polynomial data");
        }
        if (first) cols = f.polys;
        else if (f.polys != cols)
            throw Error{INSPIRE_ERR_UNSUPPORTED, "shards with different polynomial counts (resident shards share one layout)"};
        first = false;
    }
    std::stable_sort(files.begin(), files.end(), [](const File& a, const File& b) { return a.id < b.id; });
    std::vector<uint32_t> ids(files.size());
    for (size_t i = 0; i < files.size(); i++) ids[i] = files[i].id;
    Database db;
    db_begin(ctx, db, (uint32_t)files.size(), ids.data(), cols);
    ctx->dbs.erase(db_id);
    const size_t P = files.size() * (size_t)cols;
    if (P) {
        db.digits.alloc(P * ctx->ell * KN * 4);
        // slices of whole files, up to 64 MiB of u32 residues
        const size_t per_file = (size_t)cols * KN;
        const size_t files_per_slice = std::max<size_t>(1, (size_t)(16u << 20) / per_file);
        ctx->ws0.reserve(std::min(files.size(), files_per_slice) * per_file * 4);
        ctx->pin.reserve(std::min(files.size(), files_per_slice) * per_file * 4);
        std::vector<uint64_t> raw(per_file);
        for (size_t f0 = 0; f0 < files.size(); f0 += files_per_slice) {
            const size_t nf = std::min(files_per_slice, files.size() - f0);
            uint32_t* stage = ctx->pin.get<uint32_t>();
            for (size_t j = 0; j < nf; j++) {
                const File& f = files[f0 + j];
                std::ifstream in(f.path, std::ios::binary);
                in.seekg((std::streamoff)f.data_off);
                if (!in.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(per_file * 8)))
                    bad(f.path, "truncated polynomial data");
                uint32_t* dst = stage + j * per_file;
                for (size_t w = 0; w < per_file; w++) {
                    const uint64_t q = ctx->hm[(w / ctx->N) % ctx->K].q;
                    if (raw[w] >= q) bad(f.path, "a residue is not reduced");
                    dst[w] = (uint32_t)raw[w];
                }
            }
            hip_check(hipMemcpyAsync(ctx->ws0.get(), stage, nf * per_file * 4, hipMemcpyHostToDevice, ctx->stream),
                      "shard upload");
            db_digits(ctx, ctx->ws0.get<uint32_t>(), nf * cols, db.digits.get<uint32_t>() + f0 * cols * ctx->ell * KN);
            hip_check(hipStreamSynchronize(ctx->stream), "shard slice");
        }
    }
    prof_collect(ctx);
    ctx->dbs[db_id] = std::move(db);
}

// encode_database (encode_db.rs:106-164) with shards of d entries
// (setup.rs:149-153), on the device, in slices of whole shards.
void db_load_entries(inspire_ctx* ctx, uint32_t db_id, const uint8_t* data, uint64_t len, uint32_t entry_size) {
    if (entry_size == 0 || len == 0) {  // encode_db.rs:112-114: no shards
        db_load(ctx, db_id, 0, nullptr, 0, nullptr);
        return;
    }
    const uint64_t total = len / entry_size, d = ctx->N;
    const uint32_t cols = (entry_size * 8 + 15) / 16;
    const uint64_t nshards64 = (total + d - 1) / d;
    if (nshards64 > 0x7fffffffull) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "too many shards"};
    const uint32_t nshards = (uint32_t)nshards64;
    const size_t KN = (size_t)ctx->K * d;
    Database db;
    db_begin(ctx, db, nshards, nullptr, cols);
    ctx->dbs.erase(db_id);
    db.digits.alloc((size_t)nshards * cols * ctx->ell * KN * 4);
    // slices of whole shards, bounded by a 256 MiB residue staging buffer
    const uint32_t slice = (uint32_t)std::max<size_t>(1, (size_t)(256u << 20) / ((size_t)cols * KN * 4));
    ctx->ws0.reserve((size_t)std::min(slice, nshards) * cols * KN * 4);
    ctx->ws5.reserve((size_t)std::min(slice, nshards) * d * entry_size);
    for (uint32_t s0 = 0; s0 < nshards; s0 += slice) {
        const uint32_t ns = std::min(slice, nshards - s0);
        const uint64_t e0 = (uint64_t)s0 * d;
        const uint64_t ne = std::min<uint64_t>((uint64_t)ns * d, total - e0);
        hip_check(hipMemcpyAsync(ctx->ws5.get(), data + e0 * entry_size, ne * entry_size, hipMemcpyHostToDevice,
                                 ctx->stream),
                  "entries upload");
        hip_check(launch_encode_entries(ctx->ws5.get<uint8_t>(), ne, entry_size, cols, ns, ctx->ws0.get<uint32_t>(),
                                        ctx->dev, ctx->stream),
                  "encode entries");
        db_digits(ctx, ctx->ws0.get<uint32_t>(), (size_t)ns * cols,
                  db.digits.get<uint32_t>() + (size_t)s0 * cols * ctx->ell * KN);
        hip_check(hipStreamSynchronize(ctx->stream), "db slice");
    }
    prof_collect(ctx);
    ctx->dbs[db_id] = std::move(db);
}

// ---------------------------------------------------------------------------
// General external product
// ---------------------------------------------------------------------------

void external_product_batch(inspire_ctx* ctx, uint64_t count, const uint64_t* rlwe, const uint64_t* rgsw,
                            uint64_t* out) {
    if (!count) return;
    const size_t KN = (size_t)ctx->K * ctx->N, rows = 2 * ctx->ell;
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(count, 1024));
    ctx->ws0.reserve(chunk * 2 * KN * 4);
    ctx->ws1.reserve(chunk * rows * KN * 4);
    ctx->ws2.reserve(chunk * rows * 2 * KN * 4);
    uint64_t mont[kMaxK];
    for (uint32_t k = 0; k < ctx->K; k++) mont[k] = mont_of(ctx, k);
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
// Device stages
// ---------------------------------------------------------------------------

// Rows from the caller's buffers straight into the workspace: queries [0, S)
// are seeded (a expanded from seeds [S][ell][8], b read from bsrc
// [S][ell][K][N]), queries [S, Q) bring full rows fsrc [Q-S][ell][2][K][N].
// Every row is transformed (NTT, Montgomery form) from its source into
// rows_ws [Q][ell][2][K][N]; no copy of the inputs is made.
void stage_rows_from(inspire_ctx* ctx, uint32_t Q, uint32_t S, uint32_t* rows_ws, const uint32_t* seeds,
                     const uint32_t* bsrc, const uint32_t* fsrc, uint32_t* d_err, uint32_t a_res,
                     const uint32_t* a_full) {
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    uint64_t mont[kMaxK];
    for (uint32_t k = 0; k < ctx->K; k++) mont[k] = mont_of(ctx, k);
    const ScaleDev sc = scale_fwd(ctx, mont);
    if (S && ctx->logn == 11) {  // expansion and both transforms in one pass
        const size_t count = (size_t)S * ell;
        const uint32_t ka = a_res ? a_res : ctx->K;
        prof_launch(ctx, P_EXPAND, "seeded rows", [=, dev = ctx->dev, st = ctx->stream](LaunchEv ev) {
            return launch_rows_seeded(seeds, bsrc, rows_ws, count, d_err, dev, sc, st, ev, ka, a_full);
        });
    } else if (S) {
        {
            ProfScope ps(ctx, P_EXPAND);
            hip_check(launch_expand_seeds(seeds, rows_ws, (size_t)S * ell, 2 * KN, d_err, ctx->dev, ctx->stream),
                      "expand seeds");
        }
        ProfScope ps(ctx, P_ROWS_NTT);
        hip_check(launch_ntt_strided(rows_ws, 2 * KN, rows_ws, 2 * KN, (size_t)S * ell, ctx->dev, sc, false,
                                     ctx->stream),
                  "rows ntt a");
        hip_check(launch_ntt_strided(bsrc, KN, rows_ws + KN, 2 * KN, (size_t)S * ell, ctx->dev, sc, false, ctx->stream),
                  "rows ntt b");
    }
    if (Q > S) {
        ProfScope ps(ctx, P_ROWS_NTT);
        hip_check(launch_ntt_strided(fsrc, KN, rows_ws + (size_t)S * ell * 2 * KN, KN, (size_t)(Q - S) * ell * 2,
                                     ctx->dev, sc, false, ctx->stream),
                  "rows ntt full");
    }
}

// InsPIRe^0 on staged rows: out [Q][1 + cols][2][K][N] (combined, then columns).
void run_respond0(inspire_ctx* ctx, uint32_t cols, uint32_t Q, const uint32_t* rows, const uint64_t* d_qdig,
                  uint32_t* out) {
    {
        ProfScope ps(ctx, P_FIRST_DIM);
        // polys per query: 2 (combined) + 2 per column; column c at polys 2 + 2c, 3 + 2c
        hip_check(launch_first_dim(d_qdig, rows, out, Q, cols, 2 * (1 + cols), 2, ctx->dev, ctx->stream),
                  "first dim");
    }
    {
        ProfScope ps(ctx, P_COLS_INTT);
        // One launch over every poly of the response blocks.  The combined slot
        // (polys 0, 1 of each block) holds scratch here; its transform is
        // discarded because the fold below overwrites it.
        hip_check(launch_ntt(out, (size_t)Q * 2 * (1 + cols), ctx->dev, scale_inv(ctx, nullptr), true, ctx->stream),
                  "cols intt");
    }
    ProfScope ps(ctx, P_FOLD);
    hip_check(launch_sum_cols(out, Q, cols, ctx->dev, ctx->stream), "fold");
}

// ---------------------------------------------------------------------------
// Respond (host API)
// ---------------------------------------------------------------------------

const uint32_t* fetch_results(inspire_ctx* ctx, const uint32_t* d_out, size_t words, const uint32_t* d_err,
                              uint32_t* err) {
    const size_t err_at = (words * 4 + 255) & ~size_t(255);
    ctx->hout.reserve(err_at + 4);
    uint8_t* h = ctx->hout.get();
    if (words) hip_check(hipMemcpyAsync(h, d_out, words * 4, hipMemcpyDeviceToHost, ctx->stream), "results");
    hip_check(hipMemcpyAsync(h + err_at, d_err, 4, hipMemcpyDeviceToHost, ctx->stream), "error word");
    hip_check(hipStreamSynchronize(ctx->stream), "respond");
    std::memcpy(err, h + err_at, 4);
    return reinterpret_cast<const uint32_t*>(h);
}

namespace {

// Device inputs of one host-buffer group: transformed rows, the per-query
// digit addresses and the error word of the seed expansion.
struct HostRows {
    const uint32_t* rows;
    const uint64_t* qdig;
    uint32_t* err;
};

// Stages the RGSW rows ell..2ell-1 of `items` (seeded ones first, [0, S))
// from the caller's u64 buffers: narrowed into pinned memory on several
// threads, one copy to the device, then seed expansion and transforms
// (stage_rows_from).  Device layout, each part 256-byte aligned: digit
// addresses [Q] u64, error word, seeds [S][ell][8], b rows [S][ell][K][N],
// full rows [Q-S][ell][2][K][N].
HostRows upload_rows(inspire_ctx* ctx, const inspire_query* qs, const std::vector<BatchItem>& items, uint32_t S,
                     uint32_t a_res) {
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    const uint32_t Q = (uint32_t)items.size();
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_err = al((size_t)Q * 8), o_seeds = o_err + 256, o_b = al(o_seeds + (size_t)S * ell * 32),
                 o_full = al(o_b + (size_t)S * ell * KN * 4), bytes = al(o_full + (size_t)(Q - S) * ell * 2 * KN * 4);
    ctx->hin.reserve(bytes);
    ctx->ws4.reserve(bytes);
    uint8_t* h = ctx->hin.get();
    uint8_t* d = ctx->ws4.get<uint8_t>();
    uint64_t* hdig = reinterpret_cast<uint64_t*>(h);
    *reinterpret_cast<uint32_t*>(h + o_err) = 0;
    par_ranges(Q, 4, [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; t++) {
            const inspire_query& q = qs[items[t].idx];
            hdig[t] = items[t].dig;
            if (t < S) {  // seeded: rgsw [2 ell][K][N] b words, rgsw_seeds [2 ell][32]
                narrow_u64(reinterpret_cast<uint32_t*>(h + o_b) + t * ell * KN, q.rgsw + ell * KN, ell * KN);
                std::memcpy(h + o_seeds + t * ell * 32, q.rgsw_seeds + ell * 32, ell * 32);
            } else {  // full: rgsw [2 ell][2][K][N]
                narrow_u64(reinterpret_cast<uint32_t*>(h + o_full) + (t - S) * ell * 2 * KN, q.rgsw + ell * 2 * KN,
                           ell * 2 * KN);
            }
        }
    });
    hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream), "rows upload");
    ctx->ws1.reserve((size_t)Q * ell * 2 * KN * 4);
    uint32_t* rows = ctx->ws1.get<uint32_t>();
    uint32_t* d_err = reinterpret_cast<uint32_t*>(d + o_err);
    stage_rows_from(ctx, Q, S, rows, reinterpret_cast<const uint32_t*>(d + o_seeds),
                    reinterpret_cast<const uint32_t*>(d + o_b), reinterpret_cast<const uint32_t*>(d + o_full), d_err,
                    a_res);
    return HostRows{rows, reinterpret_cast<const uint64_t*>(d), d_err};
}

void fail(int32_t* status, uint32_t i, int code, const std::string& msg) {
    status[i] = code;
    set_last_error(msg);
}

}  // namespace

// The reference's checks, in its order, for one query.  Returns the database
// when the query must be computed; nullptr when it finished here (an error,
// or an empty shard answered with zeros).
static const Database* admit(inspire_ctx* ctx, const inspire_query& q, uint32_t v, inspire_response* r,
                             int32_t* status, uint32_t i, Route route, uint64_t* dig_out, uint8_t* kind) {
    const size_t KN = (size_t)ctx->K * ctx->N;
    r->num_column_ciphertexts = 0;
    r->ciphertext_b_dim = ctx->N;
    if (v > INSPIRE_VARIANT_TWO_PACKING) { fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT, "unknown variant"); return nullptr; }
    *kind = v == INSPIRE_VARIANT_NO_PACKING ? KIND_PLAIN : KIND_INSPIRING;
    if (*kind != KIND_PLAIN && route == Route::Tree) *kind = KIND_TREE;
    if (*kind != KIND_PLAIN && route == Route::ByVariant) {  // respond_with_variant (respond.rs:135-149)
        if (q.packing_mode == INSPIRE_PACKING_TREE) {
            *kind = KIND_TREE;
        } else if (!q.has_packing_keys) {
            fail(status, i, INSPIRE_ERR_PACKING_KEYS_MISSING_TREE,
                 "InspiRING packing keys missing (set packing_mode=tree to use tree packing)");
            return nullptr;
        }
    }
    const bool packed = *kind == KIND_INSPIRING;  // InspiRING's key checks (tree packing reads none)
    if (packed && !q.has_packing_keys) {  // respond.rs:296-299
        fail(status, i, INSPIRE_ERR_PACKING_KEYS_MISSING, "InspiRING client packing keys missing from query");
        return nullptr;
    }
    auto& dbs = root(ctx)->dbs;
    auto it = dbs.find(q.db_id);
    if (it == dbs.end()) {
        fail(status, i, INSPIRE_ERR_UNKNOWN_DATABASE, "unknown database " + std::to_string(q.db_id));
        return nullptr;
    }
    const Database& db = it->second;
    const int64_t slot = db.slot(q.shard_id);
    if (slot < 0) {  // respond.rs:78-82
        fail(status, i, INSPIRE_ERR_SHARD_NOT_FOUND, "Shard " + std::to_string(q.shard_id) + " not found");
        return nullptr;
    }
    if (db.cols == 0) {  // empty shard (respond.rs:84-90, 307-313): zero ciphertext, [zero]
        std::memset(r->ciphertext, 0, 2 * KN * 8);
        r->num_column_ciphertexts = 1;
        if (r->column_ciphertexts) {
            if (r->column_capacity < 1) {
                fail(status, i, INSPIRE_ERR_CAPACITY, "column_ciphertexts capacity too small");
                return nullptr;
            }
            std::memset(r->column_ciphertexts, 0, 2 * KN * 8);
        }
        return nullptr;
    }
    if (packed && q.y_all_len == 0 && !q.y_body) {  // respond.rs:365-369
        fail(status, i, INSPIRE_ERR_PACKING_KEYS_EMPTY, "InspiRING packing keys invalid: y_all and y_body are both empty");
        return nullptr;
    }
    if (*kind == KIND_PLAIN && r->column_ciphertexts && r->column_capacity < db.cols) {
        fail(status, i, INSPIRE_ERR_CAPACITY, "column_ciphertexts capacity too small");
        return nullptr;
    }
    if (*kind == KIND_TREE) {
        if (!root(ctx)->galois_count) {  // the reference indexes ServerCrs.galois_keys (automorph_pack.rs:237-250)
            fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT,
                 "tree packing needs the CRS galois keys (inspire_ctx_set_galois_keys)");
            return nullptr;
        }
        const size_t cap = r->ciphertext_capacity ? r->ciphertext_capacity : 2 * KN;
        if (cap < tree_out_words(ctx, db.cols)) {
            fail(status, i, INSPIRE_ERR_CAPACITY,
                 "ciphertext capacity too small: a tree-packed response holds " +
                     std::to_string(tree_out_words(ctx, db.cols)) + " words");
            return nullptr;
        }
    }
    *dig_out = shard_digits(db, (uint32_t)slot);
    return &db;
}

void respond_batch(inspire_ctx* ctx, uint32_t n, const inspire_query* qs, const uint32_t* variants,
                   inspire_response* rs, int32_t* status, Route route) {
    const size_t KN = (size_t)ctx->K * ctx->N;
    std::map<uint32_t, std::vector<BatchItem>> groups;  // by column count: one launch spans databases
    for (uint32_t i = 0; i < n; i++) {
        status[i] = INSPIRE_OK;
        const inspire_query& q = qs[i];
        inspire_response& r = rs[i];
        if (!r.ciphertext) { fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT, "response ciphertext buffer is null"); continue; }
        if (!q.rgsw || (q.seeded && !q.rgsw_seeds)) { fail(status, i, INSPIRE_ERR_INVALID_ARGUMENT, "query rgsw is null"); continue; }
        uint64_t dig = 0;
        uint8_t kind = KIND_PLAIN;
        const Database* db = admit(ctx, q, variants[i], &r, status, i, route, &dig, &kind);
        if (!db) continue;
        groups[db->cols].push_back(BatchItem{i, dig, kind});
    }

    for (auto& kv : groups) {
        const uint32_t cols = kv.first;
        for (uint8_t pass = 0; pass < 3; pass++) {  // KIND_PLAIN, KIND_INSPIRING, KIND_TREE
            std::vector<BatchItem> items;
            for (const BatchItem& it : kv.second)
                if (it.kind == pass) items.push_back(it);
            if (items.empty()) continue;
            // seeded queries first: their rows are expanded and transformed in one launch
            std::stable_partition(items.begin(), items.end(),
                                  [&](const BatchItem& it) { return qs[it.idx].seeded != 0; });
            const uint32_t Q = (uint32_t)items.size();
            uint32_t S = 0;
            while (S < Q && qs[items[S].idx].seeded) S++;
            const HostRows in =
                upload_rows(ctx, qs, items, S, pass == KIND_INSPIRING ? pack_rows_a_residues(ctx, cols) : 0);
            if (pass == KIND_INSPIRING) {
                pack_respond_host(ctx, cols, items.data(), Q, in.rows, in.qdig, in.err, qs, rs, status);
                continue;
            }
            const bool plain = pass == KIND_PLAIN;
            const size_t per = plain ? (size_t)(1 + cols) * 2 * KN : tree_out_words(ctx, cols);
            ctx->ws2.reserve(Q * per * 4);
            uint32_t* out = ctx->ws2.get<uint32_t>();
            if (plain) run_respond0(ctx, cols, Q, in.rows, in.qdig, out);
            else tree_run(ctx, cols, Q, in.rows, in.qdig, out);
            uint32_t herr = 0;
            const uint32_t* hout = fetch_results(ctx, out, Q * per, in.err, &herr);
            if (herr) {
                for (uint32_t t = 0; t < Q; t++)
                    fail(status, items[t].idx, INSPIRE_ERR_DEVICE, "seed expansion exhausted its draw budget");
                continue;
            }
            const uint32_t b_dim = plain ? ctx->N : tree_b_dim(ctx, cols);
            par_ranges(Q, 2, [&](size_t t0, size_t t1) {
                for (size_t t = t0; t < t1; t++) {
                    inspire_response& r = rs[items[t].idx];
                    const uint32_t* src = hout + t * per;
                    if (plain) {
                        widen_u32(r.ciphertext, src, 2 * KN);
                        r.num_column_ciphertexts = cols;
                        if (r.column_ciphertexts) widen_u32(r.column_ciphertexts, src + 2 * KN, (size_t)cols * 2 * KN);
                    } else {
                        widen_u32(r.ciphertext, src, per);
                        r.ciphertext_b_dim = b_dim;
                        r.num_column_ciphertexts = 0;  // respond.rs:262-267
                    }
                }
            });
        }
    }
    prof_collect(ctx);
}

// ---------------------------------------------------------------------------
// Respond (device-resident batch)
// ---------------------------------------------------------------------------

// The database a device batch reads, after the checks of the entry point.
static const Database& device_batch_db(inspire_ctx* ctx, const inspire_device_batch& b, bool* tree, bool* packed) {
    if (!b.shard_ids || !b.rows || !b.out || !b.status || (b.seeded && !b.seeds))
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "null argument in device batch"};
    if (b.variant > INSPIRE_VARIANT_TWO_PACKING) throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "unknown variant"};
    if (b.variant != INSPIRE_VARIANT_NO_PACKING && b.packing_mode > INSPIRE_PACKING_TREE)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "unknown packing mode"};
    *tree = b.variant != INSPIRE_VARIANT_NO_PACKING && b.packing_mode == INSPIRE_PACKING_TREE;
    *packed = b.variant != INSPIRE_VARIANT_NO_PACKING && !*tree;  // InspiRING
    if (*packed && !b.y_body)
        throw Error{INSPIRE_ERR_PACKING_KEYS_MISSING, "InspiRING client packing keys missing from query"};
    if (*tree && !root(ctx)->galois_count)
        throw Error{INSPIRE_ERR_INVALID_ARGUMENT, "tree packing needs the CRS galois keys (inspire_ctx_set_galois_keys)"};
    auto& dbs = root(ctx)->dbs;
    auto it = dbs.find(b.db_id);
    if (it == dbs.end()) throw Error{INSPIRE_ERR_UNKNOWN_DATABASE, "unknown database " + std::to_string(b.db_id)};
    const Database& db = it->second;
    if (db.cols == 0) throw Error{INSPIRE_ERR_UNSUPPORTED, "device batches need a database with columns"};
    return db;
}

DeviceBatchPlan device_batch_plan(inspire_ctx* ctx, const inspire_device_batch& b) {
    DeviceBatchPlan P{};
    const Database& db = device_batch_db(ctx, b, &P.tree, &P.packed);
    P.cols = db.cols;
    const size_t KN = (size_t)ctx->K * ctx->N, Q = b.count;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    P.o_qdig = 0;
    P.o_rows = al((Q + 1) * 8);
    P.bytes = P.o_rows + al(Q * ctx->ell * 2 * KN * 4);
    if (P.packed) {
        P.tables = find_pack_tables(ctx, db.cols);
        if (P.tables) {
            const PackWsSizes z = pack_ws_sizes(ctx, *P.tables, b.count);
            P.o_rpre = P.bytes;
            P.o_bpart = P.o_rpre + al(z.rpre * 4);
            P.o_cols = P.o_bpart + al(z.bpart * 8);
            P.o_y = P.o_cols + al(z.cols * 4);
            P.bytes = P.o_y + (pack_ybody_in_chain(ctx) ? 0 : al(Q * ctx->ell * KN * 4));
        }
    }
    return P;
}

void respond_batch_device(inspire_ctx* ctx, const inspire_device_batch& b, uint8_t* scratch, uint64_t* host_meta) {
    const size_t KN = (size_t)ctx->K * ctx->N, ell = ctx->ell;
    const uint32_t Q = b.count;
    if (!Q) return;
    bool tree = false, packed = false;
    const Database& db = device_batch_db(ctx, b, &tree, &packed);
    // per-query digit addresses and the error word, staged on the host (a
    // missing shard reads shard 0's digits and reports SHARD_NOT_FOUND)
    DeviceBatchPlan P{};
    if (scratch) P = device_batch_plan(ctx, b);
    uint64_t* d_qdig;
    uint64_t* hdig;
    uint32_t* rows;
    PackWs pws;
    uint32_t* y_ws = nullptr;
    if (scratch) {
        d_qdig = reinterpret_cast<uint64_t*>(scratch + P.o_qdig);
        rows = reinterpret_cast<uint32_t*>(scratch + P.o_rows);
        hdig = host_meta;
        if (packed) {
            pws.rpre = reinterpret_cast<uint32_t*>(scratch + P.o_rpre);
            pws.bpart = reinterpret_cast<uint64_t*>(scratch + P.o_bpart);
            pws.cols = reinterpret_cast<uint32_t*>(scratch + P.o_cols);
            y_ws = reinterpret_cast<uint32_t*>(scratch + P.o_y);
        }
    } else {
        const size_t meta = (size_t)Q * 8 + 8;
        ctx->ws4.reserve(meta + 64);
        ctx->pin.reserve(meta + 8);
        d_qdig = ctx->ws4.get<uint64_t>();
        hdig = ctx->pin.get<uint64_t>();
        ctx->ws1.reserve((size_t)Q * ell * 2 * KN * 4);
        rows = ctx->ws1.get<uint32_t>();
    }
    uint32_t* d_err = reinterpret_cast<uint32_t*>(d_qdig + Q);
    for (uint32_t t = 0; t < Q; t++) {
        const int64_t sl = db.slot(b.shard_ids[t]);
        const bool found = sl >= 0;
        hdig[t] = shard_digits(db, found ? (uint32_t)sl : 0);
        b.status[t] = found ? INSPIRE_OK : INSPIRE_ERR_SHARD_NOT_FOUND;
    }
    hdig[Q] = 0;  // the error word
    uint32_t* herr = reinterpret_cast<uint32_t*>(hdig + Q + 1);
    auto enqueue = [&] {
        hip_check(hipMemcpyAsync(d_qdig, hdig, (size_t)Q * 8 + 8, hipMemcpyHostToDevice, ctx->stream), "digits");
        // rows: transformed from the caller's buffers into the workspace layout
        if (b.seeded)
            stage_rows_from(ctx, Q, Q, rows, b.seeds, b.rows, nullptr, d_err,
                            packed ? pack_rows_a_residues(ctx, db.cols) : 0);
        else stage_rows_from(ctx, Q, 0, rows, nullptr, nullptr, b.rows, d_err);
        if (tree) tree_run(ctx, db.cols, Q, rows, d_qdig, b.out);
        else if (!packed) run_respond0(ctx, db.cols, Q, rows, d_qdig, b.out);
        else pack_run_device_ybody(ctx, db.cols, Q, rows, d_qdig, b.y_body, b.out, pws, y_ws);
        hip_check(hipMemcpyAsync(herr, d_err, 4, hipMemcpyDeviceToHost, ctx->stream), "err");
    };
    if (!scratch && ctx->graphs.enabled) {
        // every value the launches take: the batch's shape and buffers, the
        // workspaces (grown by an earlier call of this shape), the resident
        // tables they read, and which kernels are bracketed by profiling events
        const inspire_ctx* R = root(ctx);
        const std::vector<uint64_t> key = {
            Q, db.cols, (uint64_t)tree | (uint64_t)packed << 1 | (uint64_t)(b.seeded != 0) << 2,
            reinterpret_cast<uint64_t>(b.rows), reinterpret_cast<uint64_t>(b.seeds),
            reinterpret_cast<uint64_t>(b.y_body), reinterpret_cast<uint64_t>(b.out),
            reinterpret_cast<uint64_t>(ctx->ws0.get()), reinterpret_cast<uint64_t>(ctx->ws1.get()),
            reinterpret_cast<uint64_t>(ctx->ws2.get()), reinterpret_cast<uint64_t>(ctx->ws3.get()),
            reinterpret_cast<uint64_t>(ctx->ws4.get()), reinterpret_cast<uint64_t>(ctx->ws5.get()),
            reinterpret_cast<uint64_t>(ctx->ws6.get()), reinterpret_cast<uint64_t>(ctx->pin.get()),
            reinterpret_cast<uint64_t>(R->galois.get()), reinterpret_cast<uint64_t>(R->yhat.get()),
            (ctx->prof.on && !ctx->prof.suspended ? (uint64_t)ctx->prof.mask | 1ull << 32 : 0ull) |
                (chain_order_on(ctx) ? 1ull << 33 : 0ull)};
        run_with_graphs(ctx, key, enqueue);
    } else {
        enqueue();
    }
    hip_check(hipStreamSynchronize(ctx->stream), "device batch");
    // the batch's events are complete; their times are read when they are
    // queried (or once enough accumulate), not inside every batch call
    if (ctx->prof.spans.size() >= 256) prof_collect(ctx);
    if (*herr) throw Error{INSPIRE_ERR_DEVICE, "seed expansion exhausted its draw budget"};
}

}  // namespace inspire
