// Host-side context of the respond path: parameters, device tables, resident
// databases, packing tables and workspaces.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/inspire_mi355x.h"
#include "kernels.hpp"

namespace inspire {

// Error carrying an inspire status code and the reference's message text.
struct Error {
    int code;
    std::string msg;
};

void set_last_error(const std::string& msg);

// Owning device allocation.
class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t bytes);
    // Grows to at least `bytes` (contents are not preserved).
    void reserve(size_t bytes) { if (bytes > n_) { release(); alloc(bytes); } }
    void release();
    template <typename T = void>
    T* get() const { return reinterpret_cast<T*>(p_); }
    size_t bytes() const { return n_; }

   private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Owning pinned host allocation (hipHostMalloc), grown on demand.
class PinBuf {
   public:
    PinBuf() = default;
    ~PinBuf() { release(); }
    PinBuf(const PinBuf&) = delete;
    PinBuf& operator=(const PinBuf&) = delete;
    void reserve(size_t bytes);
    void release();
    template <typename T = uint8_t>
    T* get() const { return reinterpret_cast<T*>(p_); }

   private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// One resident EncodedDatabase.
struct Database {
    uint32_t num_shards = 0;
    uint32_t cols = 0;                                // polys per shard (gamma)
    std::unordered_map<uint32_t, uint32_t> slot_of;   // ShardData.id -> slot (first match)
    bool dense_ids = false;                           // ids are 0..num_shards-1 in slot order
    // slot of a shard id, or -1 (the reference's find by id, respond.rs:78-82)
    int64_t slot(uint32_t id) const {
        if (dense_ids) return id < num_shards ? (int64_t)id : -1;
        auto it = slot_of.find(id);
        return it == slot_of.end() ? -1 : (int64_t)it->second;
    }
    DevBuf digits;                                    // [shard][col][ell][K][N] u32, NTT domain
    size_t digit_words_per_shard = 0;                 // cols * ell * K * N
};

// InspiRING tables for one number of packed columns gamma (PackParams +
// OfflinePackingKeys, inspiring2.rs:107-182, 448-506), resident.
struct PackTables {
    uint32_t gamma = 0;
    DevBuf stepu;      // [gamma][K][N] u32: w_s^(g^(n-i)) (step 1 by Horner), Montgomery form;
                       // built only when the fused front end does not apply
    uint32_t ginvm[2] = {0, 0};  // gamma^-1 2^32 mod q_k
    bool odd_gens = true;      // every g^i (i < gamma - 1) odd: rotations of y_body are NTT-slot permutations
    DevBuf perm;       // [gamma][N] u32: automorphism table for g^i
    DevBuf w_all;      // [(gamma-1)][ell][K][N] u32 NTT domain, Montgomery form
    std::vector<uint32_t> gen_pows;  // g^i mod 2n, i < n
    bool fast_front = false;   // gamma = 2^k, gamma^2 | 2n: step 1 as a gamma-point DFT
    DevBuf twist;              // [K][gamma][N] gamma^-1 w_s^j, Montgomery form
    DevBuf dftw;               // [K][gamma/2][gamma/2] (W_c^m, companion), W_c = zeta^-1 of class c
};

// HIP-event timing of each kernel category on the context's stream.
enum ProfId {
    P_ROWS_NTT, P_EXPAND, P_FIRST_DIM, P_COLS_INTT, P_FOLD, P_B0, P_STEP1, P_Y_NTT, P_CHAIN, P_DB, P_FRONT,
    P_TREE_LEAVES, P_TREE_COMBINE, P_EXTRACT, P_COUNT
};
extern const char* const kProfNames[P_COUNT];

struct Profiler {
    bool on = false;
    bool suspended = false;  // no events recorded (the no-allocation entry point)
    uint32_t mask = ~0u;  // kernels (bit = ProfId) bracketed while on
    struct Span { int id; hipEvent_t a, b; };
    std::vector<Span> spans;
    std::vector<hipEvent_t> pool;
    double total_ms[P_COUNT] = {};
    uint64_t launches[P_COUNT] = {};
};

// Client packing keys registered once and kept resident: the NTT of y_body
// (Montgomery form), [slot][ell][K][N] u32.
struct ClientPool {
    DevBuf keys;
    uint32_t capacity = 0, used = 0;
    std::unordered_map<uint32_t, uint32_t> slot_of;  // client id -> slot
};

// A device batch's launch sequence captured once and replayed (hipGraph):
// the key is every value its launches take as arguments (batch shape,
// caller buffers, workspaces, staging, resident tables) and which kernels
// are profiled.  Events recorded inside a graph do not time its kernels on
// this runtime, so a profiled kernel splits the sequence: a launch that
// records its own events (prof_launch) is replayed eagerly between the
// graph segments before and after it, a profiled scope (ProfScope) is a
// graph segment of its own between two events recorded on the stream.
// Unprofiled, the whole sequence is one graph.
// A kernel launch given the events it records (null: none).
using LaunchFn = std::function<hipError_t(LaunchEv)>;
struct GraphSeg {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int prof = -1;  // the ProfId this segment's events time, or -1
    // a profiled launch that records its own events: kept out of the graphs
    // and launched between them at every replay (graph and exec are null)
    LaunchFn eager;
    const char* what = nullptr;
};
struct GraphEntry {
    std::vector<uint64_t> key;
    std::vector<GraphSeg> segs;
};
struct GraphCache {
    bool enabled = true;
    std::vector<std::vector<uint64_t>> last_eager;  // keys of the last calls run eagerly
    std::vector<GraphEntry> entries;    // least recently used first
    uint64_t replays = 0, captures = 0;
    // while a profiled sequence is captured: the segments closed so far, and
    // the first error of a segment boundary (profiling scopes end in
    // destructors, which report here instead of throwing)
    bool segmenting = false;
    std::vector<GraphSeg> pending;
    hipError_t seg_error = hipSuccess;
};
// Ends the capture of the current segment and begins the next: the closed
// segment is kept (with `prof`, the ProfId it holds, or -1) when it has nodes.
void graph_segment(inspire_ctx* ctx, int prof) noexcept;
void graph_cache_clear(GraphCache& G);

struct HostModulus {
    uint64_t q = 0, psi = 0;
    std::vector<uint32_t> tw, twp, itw, itwp;
};

}  // namespace inspire

struct inspire_ctx {
    inspire_params params{};
    uint32_t N = 0, K = 0, logn = 0, ell = 0;
    uint64_t q = 0, delta = 0;
    uint8_t w_seed[32]{};
    inspire::HostModulus hm[inspire::kMaxK];
    int device = 0;           // the HIP device current at creation (owns the stream and tables)
    inspire::DevBuf tables;   // twiddles of every modulus
    inspire::DevBuf mfma_tables;  // matrix-core NTT constants (d = 2048, moduli below 2^28)
    inspire::CtxDev dev{};
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::map<uint32_t, inspire::Database> dbs;
    std::map<uint32_t, std::unique_ptr<inspire::PackTables>> packs;
    inspire::ClientPool clients;
    // ServerCrs.galois_keys (tree packing): [log N][ell][2][K][N] NTT domain,
    // Montgomery form, and NTT(X^(N >> l)) per level, [log N][K][N]
    inspire::DevBuf galois, yhat;
    uint32_t galois_count = 0;
    // Writers (database load/unload, client registration) hold it exclusively,
    // serving lanes share it while a batch reads databases, tables and keys.
    std::shared_mutex rw;
    std::mutex packs_mu;  // the packing-table map (built on first use, never erased)
    // A serving lane: its own stream and workspaces; databases, tables and
    // client keys belong to the owner context.
    inspire_ctx* owner = nullptr;
    // scratch reused across calls
    inspire::DevBuf ws0, ws1, ws2, ws3, ws4, ws5, ws6;
    inspire::PinBuf pin;  // small per-batch metadata
    // pinned staging of the host-buffer calls: inputs (rows, seeds, digit
    // addresses), client keys, responses
    inspire::PinBuf hin, hkeys, hout;
    std::vector<uint32_t> hbuf;
    inspire::Profiler prof;
    inspire::GraphCache graphs;
    // Packing chains of an owner and its lanes submitted within chain_window_us
    // of each other run one after another: the later launch waits on the
    // event recorded after the earlier one, whichever stream ran it.
    // chain_done is this context's; the owner keeps the last one recorded and
    // when its chain was submitted.
    hipEvent_t chain_done = nullptr;
    std::mutex chain_mu;  // owner: guards last_chain, last_submit
    hipEvent_t last_chain = nullptr;
    std::chrono::steady_clock::time_point last_submit{};
    std::atomic<uint32_t> chain_window_us{60};  // owner: 0 turns the ordering off
};

namespace inspire {
inline inspire_ctx* root(inspire_ctx* c) { return c->owner ? c->owner : c; }
inline const inspire_ctx* root(const inspire_ctx* c) { return c->owner ? c->owner : c; }

// Records HIP events around the launches in its scope when profiling is on.
// While a profiled sequence is captured into a graph, the scope is a segment
// of its own instead (graph_segment at both ends) and records no event.
class ProfScope {
   public:
    ProfScope(inspire_ctx* ctx, int id);
    ~ProfScope();

   private:
    inspire_ctx* ctx_;
    int id_;
    hipEvent_t a_ = nullptr;
    bool seg_ = false;
};
// One kernel launch that records its own events (hipExtLaunchKernelGGL):
// `launch` enqueues the kernel with the events it is given, null when the
// kernel is not profiled.  A profiled span is the kernel's execution,
// without the host's submission time that a scope around a launch on an idle
// stream includes.  While a profiled sequence is captured into graphs, the
// launch is not captured: it is kept (with the values it captured) to run
// between the graph segments at every replay, with fresh events.
// ordered: a packing-chain launch, run after a chain of the owner's other
// lanes submitted just before it (inspire_ctx::chain_window_us); while a
// sequence is captured into graphs it is kept out of them like a profiled
// launch.
void prof_launch(inspire_ctx* ctx, int id, const char* what, LaunchFn launch, bool ordered = false);
// The chain ordering applies to this context's calls.
bool chain_order_on(const inspire_ctx* ctx);
// Folds finished spans into the totals (synchronises the stream).
void prof_collect(inspire_ctx* ctx);

// Shared hold on the owner's resident state for a call on a lane (the owner's
// own calls are serialised by its mutex; writers hold the lock exclusively).
struct LaneRead {
    std::shared_lock<std::shared_mutex> lk;
    explicit LaneRead(inspire_ctx* c) {
        if (c->owner) lk = std::shared_lock<std::shared_mutex>(c->owner->rw);
    }
};

// Records no profiling events while in scope.
struct ProfSuspend {
    inspire_ctx* ctx;
    bool was;
    explicit ProfSuspend(inspire_ctx* c) : ctx(c), was(c->prof.suspended) { c->prof.suspended = true; }
    ~ProfSuspend() { ctx->prof.suspended = was; }
};

// Looks up the no-allocation entry point's device scratch on `device` (the
// current device) once per device; called when a context is created.
void noalloc_scratch_setup(int device);

// A lane of `owner`: its own stream and workspaces over the owner's resident
// state (databases, tables, keys).
inspire_ctx* make_lane_ctx(inspire_ctx* owner);
void destroy_lane_ctx(inspire_ctx* lane);
}  // namespace inspire
