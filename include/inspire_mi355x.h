/*
 * inspire_mi355x.h -- C ABI of the MI355X (gfx950) InsPIRe server respond path.
 *
 * This is the drop-in boundary for the respond path of igor53627/inspire-rs
 * (Rust crate `inspire` 0.1.2).  Each entry point replaces one public Rust
 * function; the replaced function is cited as file:line in the reference.  A
 * Rust caller binds these with `extern "C"` declarations of the same
 * signatures and #[repr(C)] mirrors of the structs below.
 *
 * Conventions
 *  - Polynomials cross the boundary as the reference's `Poly.coeffs`
 *    (src/math/poly.rs:63-77): uint64_t[K][d], residue-major, canonical
 *    residues in [0, q_k).  An RLWE ciphertext is (a, b) = uint64_t[2][K][d];
 *    an RGSW ciphertext is its 2*ell rows, uint64_t[2*ell][2][K][d]
 *    (src/rgsw/types.rs:104-110).
 *  - NTT-domain values crossing the boundary are in the reference's
 *    representation (Montgomery form with R = 2^64, src/math/ntt.rs:223-246).
 *  - Every call returns INSPIRE_OK (0) or an error code; the message text of
 *    the last error on the calling thread is inspire_last_error().  Error
 *    codes that mirror the reference's PirError strings (src/pir/error.rs)
 *    carry the same text.
 *  - All inputs are borrowed and read-only; outputs are caller-owned buffers.
 *    A context may be shared by several threads: calls are serialised inside.
 *  - The context keeps the preprocessed databases and the CRS-derived
 *    packing tables resident in device memory (HBM).
 */
#ifndef INSPIRE_MI355X_H
#define INSPIRE_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define INSPIRE_OK 0
#define INSPIRE_ERR_INVALID_ARGUMENT 1
#define INSPIRE_ERR_INVALID_PARAMS 2       /* InspireParams::validate (src/params.rs:348-387) */
#define INSPIRE_ERR_UNSUPPORTED 3          /* valid for the reference, not on this path (message says why) */
#define INSPIRE_ERR_SHARD_NOT_FOUND 4      /* "Shard {id} not found" (src/pir/respond.rs:82) */
#define INSPIRE_ERR_PACKING_KEYS_MISSING 5 /* "InspiRING client packing keys missing from query" (respond.rs:299) */
#define INSPIRE_ERR_PACKING_KEYS_EMPTY 6   /* "InspiRING packing keys invalid: y_all and y_body are both empty" (respond.rs:367-369) */
#define INSPIRE_ERR_PACKING_KEYS_MISSING_TREE 7 /* "InspiRING packing keys missing (set packing_mode=tree to use tree packing)" (respond.rs:141-143) */
#define INSPIRE_ERR_UNKNOWN_DATABASE 8
#define INSPIRE_ERR_OUT_OF_MEMORY 9
#define INSPIRE_ERR_DEVICE 10
#define INSPIRE_ERR_CAPACITY 11            /* an output buffer is too small */
#define INSPIRE_ERR_PENDING 12             /* a serving-loop ticket is not finished yet */

/* InspireVariant (src/params.rs:78-103) */
#define INSPIRE_VARIANT_NO_PACKING 0
#define INSPIRE_VARIANT_ONE_PACKING 1
#define INSPIRE_VARIANT_TWO_PACKING 2

/* PackingMode (src/pir/query.rs:27-40) */
#define INSPIRE_PACKING_INSPIRING 0
#define INSPIRE_PACKING_TREE 1

/* inspire_params.flags */
/* Apply NTT-domain automorphisms to every residue (a corrected mode).  The
 * default reproduces the reference bit for bit, including its rebuild of
 * residue 1 from residue 0 in apply_automorphism_ntt
 * (src/inspiring/inspiring2.rs:279-296). */
#define INSPIRE_FLAG_RESIDUE_WISE_AUTOMORPHISM 0x1u

typedef struct inspire_ctx inspire_ctx;

/* InspireParams (src/params.rs:134-185): the fields the respond path reads.
 * sigma and security_level only matter to setup and clients. */
typedef struct {
    uint32_t ring_dim;       /* d: 256, 512, 1024, 2048 or 4096 on this path */
    uint32_t crt_count;      /* 1 or 2; each modulus below 2^29 on this path */
    uint64_t crt_moduli[2];
    uint64_t q;              /* product of crt_moduli */
    uint64_t p;              /* plaintext modulus */
    uint64_t gadget_base;    /* power of two, at most 2^32 on this path */
    uint32_t gadget_len;     /* ell */
    uint32_t flags;          /* INSPIRE_FLAG_* */
} inspire_params;

/* Query: ClientQuery / SeededClientQuery (src/pir/query.rs:119-150). */
typedef struct {
    uint32_t db_id;          /* which resident EncodedDatabase */
    uint32_t shard_id;       /* ClientQuery.shard_id */
    uint32_t seeded;         /* 0: ClientQuery, 1: SeededClientQuery */
    uint32_t packing_mode;   /* INSPIRE_PACKING_* */
    /* seeded == 0: rgsw = RgswCiphertext rows, uint64_t[2*ell][2][K][d].
     * seeded == 1: rgsw = the b polys of the SeededRlweCiphertext rows,
     *              uint64_t[2*ell][K][d], and rgsw_seeds = uint8_t[2*ell][32]. */
    const uint64_t *rgsw;
    const uint8_t *rgsw_seeds;
    /* inspiring_packing_keys: Option<ClientPackingKeys> (inspiring2.rs:603-628) */
    uint32_t has_packing_keys;
    uint32_t y_all_len;      /* number of y_all rotations (0 = empty) */
    const uint64_t *y_body;  /* uint64_t[ell][K][d], or NULL when empty */
    const uint64_t *y_all;   /* uint64_t[y_all_len][ell][K][d], or NULL */
} inspire_query;

/* ServerResponse (src/pir/respond.rs:30-37).  The ciphertext is a.coeffs
 * followed by b.coeffs: a = uint64_t[K][d], b = uint64_t[K][ciphertext_b_dim].
 * b has d coefficients except in a tree-packed response (respond_one_packing)
 * with two moduli and at least two columns, where the reference rebuilds b
 * from its 2d residue words as 2d composite coefficients
 * (src/inspiring/automorph_pack.rs:476-484, Poly::from_coeffs_moduli): then
 * ciphertext_b_dim = 2d and the ciphertext holds 3*K*d words. */
typedef struct {
    uint64_t *ciphertext;          /* out: a then b, see above */
    uint64_t *column_ciphertexts;  /* out: uint64_t[column_capacity][2][K][d], may be NULL */
    uint32_t column_capacity;
    uint32_t num_column_ciphertexts; /* out: the reference's column_ciphertexts.len() */
    uint32_t ciphertext_capacity;  /* in: words the ciphertext buffer holds; 0 means 2*K*d */
    uint32_t ciphertext_b_dim;     /* out: coefficients per residue of b (d or 2d) */
} inspire_response;

/* ---- context ---- */

/* Validates the parameters (params.rs:348-387), builds the NTT tables for the
 * reference's roots (ntt.rs:116-164) and the InspiRING tables derived from the
 * CRS seed w_seed (ServerCrs.inspiring_w_seed, src/pir/setup.rs:197-201). */
int inspire_ctx_create(const inspire_params *params, const uint8_t w_seed[32], inspire_ctx **out);
/* ServerCrs.galois_keys (src/pir/setup.rs:167-173), the key-switching
 * matrices tree packing uses: count = log2(d) matrices, matrix i for the
 * automorphism t = d/2^i + 1 (ks/setup.rs:218-247), each gadget_len rows
 * (a, b): keys = uint64_t[count][gadget_len][2][K][d], canonical residues.
 * They stay resident in device memory; setting them again replaces them. */
int inspire_ctx_set_galois_keys(inspire_ctx *ctx, const uint64_t *keys, uint32_t count);
/* A lane of ctx: a second handle on the same resident state (databases,
 * CRS tables, galois keys, registered client keys) with its own HIP stream
 * and workspaces.  Calls on different lanes, from different threads, run
 * concurrently on the device, so one lane's host work and transfers overlap
 * another's kernels (the reference's server answers concurrent requests from
 * tokio handlers over one shared state, src/bin/server.rs:53-57,120-236).
 * Loads, unloads and key registration go through the owner.  Destroy lanes
 * (inspire_ctx_destroy) before their owner. */
int inspire_ctx_create_lane(inspire_ctx *ctx, inspire_ctx **out);
void inspire_ctx_destroy(inspire_ctx *ctx);
const char *inspire_last_error(void);
/* Device memory held by the context, in bytes. */
uint64_t inspire_ctx_device_bytes(const inspire_ctx *ctx);

/* ---- resident databases: EncodedDatabase / ShardData (src/pir/setup.rs:105-120) ---- */

/* Loads shards given as encoded polynomials: polys = uint64_t[num_shards]
 * [polys_per_shard][K][d] (ShardData.polynomials), shard_ids = ShardData.id.
 * polys_per_shard == 0 loads empty shards.  Replaces a database with the same
 * db_id.  The preprocessing (Delta scaling, gadget digits, forward NTT) runs on
 * the device; the digits stay resident. */
int inspire_db_load(inspire_ctx *ctx, uint32_t db_id, uint32_t num_shards, const uint32_t *shard_ids,
                    uint32_t polys_per_shard, const uint64_t *polys);
/* encode_database (src/pir/encode_db.rs:106-164) of raw entry bytes followed
 * by inspire_db_load; shards hold d entries (setup.rs:149-153). */
int inspire_db_load_entries(inspire_ctx *ctx, uint32_t db_id, const uint8_t *data, uint64_t len,
                            uint32_t entry_size);
/* load_shard_binary (src/pir/mmap.rs:55-85) of every shard-*.bin file in
 * dir, the files save_shards_binary writes (mmap.rs:19-52): the shards are
 * ingested into device memory once, where the reference's respond_mmap*
 * (src/pir/respond.rs:492-640) read one from disk per query.  Every file must
 * match the context's ring dimension and moduli and hold the same number of
 * polynomials. */
int inspire_db_load_shard_files(inspire_ctx *ctx, uint32_t db_id, const char *dir);
int inspire_db_unload(inspire_ctx *ctx, uint32_t db_id);

/* ---- respond (src/pir/respond.rs) ---- */

/* respond (respond.rs:71-115): InsPIRe^0, full query. */
int inspire_respond(inspire_ctx *ctx, const inspire_query *q, inspire_response *r);
/* respond_seeded (respond.rs:408-415). */
int inspire_respond_seeded(inspire_ctx *ctx, const inspire_query *q, inspire_response *r);
/* respond_inspiring (respond.rs:284-386): InspiRING packed response. */
int inspire_respond_inspiring(inspire_ctx *ctx, const inspire_query *q, inspire_response *r);
/* respond_seeded_inspiring (respond.rs:391-398). */
int inspire_respond_seeded_inspiring(inspire_ctx *ctx, const inspire_query *q, inspire_response *r);
/* respond_with_variant (respond.rs:129-151) and respond_seeded_with_variant
 * (respond.rs:161-183); q->seeded selects between them. */
int inspire_respond_with_variant(inspire_ctx *ctx, const inspire_query *q, uint32_t variant,
                                 inspire_response *r);
/* respond_one_packing (respond.rs:213-268): InsPIRe^1 with tree packing of
 * the column LWEs (pack_lwes, src/inspiring/automorph_pack.rs:418-487) under
 * the CRS galois keys (inspire_ctx_set_galois_keys).  The response's b has 2d
 * coefficients per residue with two moduli and at least two columns (see
 * inspire_response); q->packing_mode is not read. */
int inspire_respond_one_packing(inspire_ctx *ctx, const inspire_query *q, inspire_response *r);
/* respond_seeded_packed (respond.rs:428-435): the same for a SeededClientQuery. */
int inspire_respond_seeded_packed(inspire_ctx *ctx, const inspire_query *q, inspire_response *r);
/* Many queries (any variants, databases and shards) in one batched launch
 * sequence.  status[i] receives each query's code; the return value is
 * INSPIRE_OK when the batch ran (individual queries may still have failed). */
int inspire_respond_batch(inspire_ctx *ctx, uint32_t n, const inspire_query *qs, const uint32_t *variants,
                          inspire_response *rs, int32_t *status);

/* ---- device-resident batches ---- */

/* A batch whose inputs already sit in device memory (HBM), all of one
 * variant and one database; the semantics are those of respond_with_variant
 * / respond_seeded_with_variant (src/pir/respond.rs:129-183) per query, with
 * InspiRING keys given as y_body (the wire form, inspiring2.rs:603-628).
 * Device layouts are u32 canonical residues:
 *   rows   seeded: b polys of RGSW rows ell..2ell-1, [count][ell][K][d]
 *          full:   rows ell..2ell-1 (a, b),           [count][ell][2][K][d]
 *          (rows 0..ell-1 multiply the zero a-digits of the database's
 *          trivial encryption and do not affect the response)
 *   seeds  seeded: [count][ell][8] (32-byte seeds of rows ell..2ell-1)
 *   y_body packed variants with InspiRING packing: [count][ell][K][d]
 *          (not read with tree packing)
 *   out    InspiRING packing: [count][2][K][d];
 *          tree packing: [count][K*d + K*b_dim] (a, then b; b_dim as in
 *          inspire_response);
 *          InsPIRe^0: [count][1 + cols][2][K][d] (ciphertext, then columns)
 * shard_ids and status are host arrays.  status[i] is INSPIRE_OK or
 * INSPIRE_ERR_SHARD_NOT_FOUND.  The call returns when the batch has finished. */
typedef struct {
    uint32_t count;
    uint32_t db_id;
    uint32_t variant;
    uint32_t seeded;
    const uint32_t *shard_ids;
    const uint32_t *rows;
    const uint32_t *seeds;
    const uint32_t *y_body;
    uint32_t *out;
    int32_t *status;
    uint32_t packing_mode;   /* INSPIRE_PACKING_* for the packed variants */
    uint32_t reserved;
} inspire_device_batch;

int inspire_respond_batch_device(inspire_ctx *ctx, const inspire_device_batch *batch);
/* Graph replay of device batches (on by default).  The launch sequence of
 * inspire_respond_batch_device is captured into a HIP graph the second time
 * the same batch shape and buffers (device pointers, count, database column
 * count, variant) arrive within eight calls, and replayed whenever they
 * recur (up to eight graphs per context or lane, least recently used
 * dropped first), so a
 * serving loop that reuses its buffers submits one graph per batch instead
 * of each kernel and copy.  The buffers' contents may change between calls;
 * results are those of the eager launches.  While profiling is enabled
 * (inspire_ctx_set_profiling), the profiled kernels split the captured
 * sequence: a kernel whose events its own dispatch records is launched
 * between the graphs at every replay, other profiled kernels are graphs of
 * their own between two HIP events recorded on the stream (the profiled set
 * is part of the batch's key).  enable = 0 runs every batch eagerly and
 * drops the captured graphs.  Neither call has a counterpart in
 * the reference (which has no device). */
int inspire_ctx_set_graphs(inspire_ctx *ctx, int enable);
/* Packing chains of lanes submitted together (set on the owner context;
 * default 60 us).  The packing chain of an InspiRING batch fills every
 * compute unit with one workgroup per query.  Lanes whose batches arrive at
 * the same moment run their chains at the same time, each beside the
 * others', then their row expansions and front ends at the same time, and
 * tend to stay in step batch after batch.  A chain launch submitted less
 * than window_us after the previous one, from another lane of the same
 * owner, therefore waits for it on the GPU (a stream wait on an event, no
 * host synchronisation), so that the lanes fall out of step.  Chains
 * submitted further apart are not ordered.  window_us = 0 turns the
 * ordering off.  No counterpart in the reference. */
int inspire_ctx_set_chain_window(inspire_ctx *ctx, uint32_t window_us);
/* The owner's current window (a lane reports its owner's). */
int inspire_ctx_chain_window(const inspire_ctx *ctx, uint32_t *window_us);
int inspire_ctx_graph_stats(const inspire_ctx *ctx, uint64_t *captures, uint64_t *replays);

/* inspire_respond_batch_device without any allocation: no hipMalloc,
 * hipFree or pinned host allocation in the call.  This entry point has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its device scratch is one 24 MiB array
 * per GPU, held by the library until the process ends: enough for 64
 * InsPIRe^2 queries at d = 2048 with 16 columns, or 256 InsPIRe^0 queries;
 * a batch needing more fails with INSPIRE_ERR_CAPACITY and the size in the
 * message, and more than 1024 queries are refused the same way.  Tree
 * packing is refused (INSPIRE_ERR_UNSUPPORTED), and so is an InspiRING batch
 * whose column count has no tables yet (any allocating respond call on the
 * context builds them).  Calls on one GPU share the scratch: they are run one
 * at a time (the call returns when its batch has finished).  No profiling
 * events are recorded. */
int inspire_respond_batch_device_noalloc(inspire_ctx *ctx, const inspire_device_batch *batch);

/* ---- serving loop ---- */

/* Registers a client's InspiRING packing keys once: y_body of its
 * ClientPackingKeys (src/inspiring/inspiring2.rs:603-628), uint64_t[ell][K][d].
 * The keys stay resident in device memory (their NTT); requests to the
 * serving loop name the client instead of carrying the keys.  Registering an
 * id again replaces its keys. */
int inspire_register_client_keys(inspire_ctx *ctx, uint32_t client_id, const uint64_t *y_body);

typedef struct inspire_server inspire_server;

typedef struct {
    uint32_t lanes;        /* batches in flight on the device at once (one HIP stream each), 1..8 */
    uint32_t max_batch;    /* queries per batch, 1..1024 */
    uint64_t hbm_budget;   /* device bytes all in-flight queries may hold; 0: lanes * max_batch
                              of the largest query footprint */
    uint32_t max_wait_us;  /* a lane with fewer than max_batch queries waits at most this long for more
                              while another lane is busy; 0: start at once */
    uint32_t reserved;
} inspire_server_config;

/* One query to the serving loop (the respond_with_variant /
 * respond_seeded_with_variant semantics of src/pir/respond.rs:129-183 with
 * InspiRING packing), inputs in device-ready form, copied at submit:
 *   seeded: rows = b polys of RGSW rows ell..2ell-1, uint32_t[ell][K][d];
 *           seeds = their 32-byte seeds, uint8_t[ell][32]
 *   full:   rows = RGSW rows ell..2ell-1, uint32_t[ell][2][K][d]
 * (rows 0..ell-1 multiply the zero a-digits of the database and do not
 * affect the response), unless flags has INSPIRE_REQUEST_BORROW.  Packed
 * variants use the registered keys of client_id.  arrival_ns: the query's
 * arrival on CLOCK_MONOTONIC (0: now); latencies are measured from it. */
typedef struct {
    uint32_t db_id, shard_id, variant, seeded, client_id;
    uint32_t flags;        /* INSPIRE_REQUEST_* */
    const uint32_t *rows;
    const uint8_t *seeds;
    uint64_t arrival_ns;
} inspire_request;

/* inspire_request.flags: the caller keeps rows and seeds unchanged until the
 * ticket's wait returns; the loop reads them in place instead of copying at
 * submit. */
#define INSPIRE_REQUEST_BORROW 0x1u

typedef struct {
    uint64_t requests;          /* completed */
    uint64_t batches;           /* launched */
    uint64_t max_batch_seen;
    uint64_t peak_inflight_bytes;
    uint64_t peak_lanes_busy;   /* most batches on the device at once */
    uint64_t launch_groups;     /* launch sequences the batches ran: one per column count
                                   and batch, InsPIRe^0 sharing the packed variants' when the
                                   fused front end applies (gamma = 2^k, gamma^2 | 2d, two moduli) */
} inspire_server_stats;

/* Starts the loop: every lane has a worker thread that takes the queries
 * waiting when it is free (up to max_batch, within the HBM budget), runs
 * them as one batch (any databases and variants together), and completes
 * them.  Queries arriving while other batches run are admitted by the next
 * free lane, so batches overlap on the device. */
int inspire_server_create(inspire_ctx *ctx, const inspire_server_config *cfg, inspire_server **out);
/* Copies n requests in and returns their tickets. */
int inspire_server_submit(inspire_server *s, uint32_t n, const inspire_request *reqs, uint64_t *tickets);
/* Waits for a ticket (timeout_ms < 0: no limit) and releases it.  r receives
 * the ServerResponse (as inspire_respond*), status the query's code, done_ns
 * its completion time on CLOCK_MONOTONIC.  Returns INSPIRE_ERR_PENDING when
 * the time ran out (the ticket stays valid). */
int inspire_server_wait(inspire_server *s, uint64_t ticket, int32_t timeout_ms, inspire_response *r,
                        int32_t *status, uint64_t *done_ns);
/* inspire_server_wait for n tickets in turn: rs[i] (may be NULL: status
 * only), status[i] and done_ns[i] (may be NULL) receive ticket i's results,
 * and status[i] is INSPIRE_ERR_PENDING for a ticket whose wait timed out
 * (it stays valid; the call then returns INSPIRE_ERR_PENDING).  One call
 * drains a run of completions, as a server's completion queue does. */
int inspire_server_wait_many(inspire_server *s, uint32_t n, const uint64_t *tickets, int32_t timeout_ms,
                             inspire_response *rs, int32_t *status, uint64_t *done_ns);
int inspire_server_get_stats(inspire_server *s, inspire_server_stats *out);
/* Finishes the queries already submitted, then stops the workers. */
void inspire_server_destroy(inspire_server *s);

/* ---- wire formats ---- */

/* The byte formats the reference's server and clients exchange, for the
 * parameter set `params` (no device is touched):
 *  - bincode 1.3 (bincode::serialize; ServerResponse::to_binary /
 *    from_binary, src/pir/respond.rs:43-51): little-endian, u64 lengths,
 *    fields in declaration order, skip_serializing_if fields left out;
 *  - serde_json, the HTTP bodies of src/bin/server.rs:120-236.
 * Polynomials are the reference's Poly (src/math/poly.rs:63-77): coeffs
 * (residue-major words), moduli, q, dim, crt_q0_inv_mod_q1, is_ntt.
 * The encoders write *len bytes; with out == NULL or a capacity below the
 * size they return INSPIRE_ERR_CAPACITY with *len set, so a first call with
 * out == NULL sizes the buffer. */
int inspire_response_to_binary(const inspire_params *params, const inspire_response *r, uint8_t *out,
                               uint64_t capacity, uint64_t *len);
/* serde_json::to_vec of the ServerResponse (the "response" member of the
 * server's QueryResponse, src/bin/server.rs:88-92). */
int inspire_response_to_json(const inspire_params *params, const inspire_response *r, char *out,
                             uint64_t capacity, uint64_t *len);
/* ServerResponse::from_binary into caller buffers (ciphertext_capacity and
 * column_capacity as for a respond call; ciphertext_b_dim is set). */
int inspire_response_from_binary(const inspire_params *params, const uint8_t *in, uint64_t len,
                                 inspire_response *r);

/* A decoded query: owns the polynomials the inspire_query filled by the
 * decoders points into.  Free it after the respond call. */
typedef struct inspire_wire_query inspire_wire_query;
/* A ClientQuery or SeededClientQuery (src/pir/query.rs:119-150) as the
 * server's /query and /query_seeded handlers receive it (serde_json; a
 * seeded query is recognised by its rows' seeds).  Fills q (db_id = 0) with
 * pointers into *out.  The polynomials must carry the parameters' moduli and
 * ring dimension, and the gadget their base and length; full-packing keys
 * (z_body) are refused with INSPIRE_ERR_UNSUPPORTED. */
int inspire_query_from_json(const inspire_params *params, const char *json, uint64_t len,
                            inspire_wire_query **out, inspire_query *q);
/* The same from bincode::serialize of a ClientQuery (seeded = 0) or a
 * SeededClientQuery (seeded = 1).  inspiring_packing_keys is absent when
 * None and ClientPackingKeys.z_body when empty (skip_serializing_if). */
int inspire_query_from_binary(const inspire_params *params, const uint8_t *in, uint64_t len, uint32_t seeded,
                              inspire_wire_query **out, inspire_query *q);
void inspire_wire_query_free(inspire_wire_query *w);

/* ---- client-side extract (src/pir/extract.rs) ---- */

/* Decrypts the coefficients an entry needs under the client's RLWE secret key
 * (RlweCiphertext::decrypt, src/rlwe/enc.rs:90-108: round((a s + b) / Delta)
 * mod p) and rebuilds the entry's bytes (reconstruct_entry,
 * src/pir/encode_db.rs:186-201), 2 bytes per column value.  secret_key =
 * ClientState.rlwe_secret_key, uint64_t[K][d] canonical residues; entry =
 * entry_size bytes out.  The response is read as inspire_response describes
 * it (ciphertext_b_dim 0 means d). */
/* extract (extract.rs:31-68): coefficient 0 of each of the first
 * ceil(entry_size / 2) column ciphertexts; a response without column
 * ciphertexts uses the ciphertext's coefficient 0 for every column. */
int inspire_extract(inspire_ctx *ctx, const inspire_response *r, const uint64_t *secret_key, uint32_t entry_size,
                    uint8_t *entry);
/* extract_with_variant (extract.rs:78-92): INSPIRE_VARIANT_NO_PACKING is
 * extract; ONE_PACKING and TWO_PACKING are extract_packed (:119-157),
 * coefficient c times d^-1 mod p, or times 1 when d has no inverse mod p. */
int inspire_extract_with_variant(inspire_ctx *ctx, const inspire_response *r, const uint64_t *secret_key,
                                 uint32_t entry_size, uint32_t variant, uint8_t *entry);
/* extract_inspiring (extract.rs:163-192): coefficient c, unscaled. */
int inspire_extract_inspiring(inspire_ctx *ctx, const inspire_response *r, const uint64_t *secret_key,
                              uint32_t entry_size, uint8_t *entry);

/* The same for a batch of responses already in device memory, one mode for
 * the batch.  Device layouts are u32 canonical residues, response i at
 * responses + i * stride words, in the layouts inspire_device_batch.out
 * gives them:
 *   a [K][d], b [K][b_dim], then num_column_ciphertexts column ciphertexts
 *   [2][K][d] (InsPIRe^0: the ciphertext, then its columns).
 * secret_keys [count][K][d]; entries [count][entry_size] bytes out.  In the
 * packed modes entry_size may need at most d coefficients (2d bytes): the
 * reference panics beyond. */
#define INSPIRE_EXTRACT_COLUMNS 0    /* extract */
#define INSPIRE_EXTRACT_PACKED 1     /* extract_packed (extract_with_variant, ONE/TWO_PACKING) */
#define INSPIRE_EXTRACT_INSPIRING 2  /* extract_inspiring */
typedef struct {
    uint32_t count;
    uint32_t mode;                    /* INSPIRE_EXTRACT_* */
    uint32_t entry_size;              /* bytes per entry */
    uint32_t num_column_ciphertexts;  /* after the ciphertext; read in INSPIRE_EXTRACT_COLUMNS mode */
    uint32_t b_dim;                   /* coefficients per residue of the ciphertext's b: d or 2d (0 means d) */
    uint32_t reserved;
    uint64_t stride;                  /* words from one response to the next; 0: K d + K b_dim +
                                         num_column_ciphertexts 2 K d */
    const uint32_t *responses;
    const uint32_t *secret_keys;
    uint8_t *entries;
} inspire_extract_batch;

/* Returns when the batch has finished. */
int inspire_extract_batch_device(inspire_ctx *ctx, const inspire_extract_batch *batch);

/* ---- profiling ---- */

/* When enabled, every kernel launch of the respond path is bracketed by HIP
 * events on the context's stream.  inspire_ctx_profile reads the summed
 * duration (ms) and launch count of one kernel by name (rows_ntt,
 * expand_seeds, first_dim, cols_intt, fold, pack_front, pack_b0, pack_step1,
 * y_ntt, pack_chain, db_load, tree_leaves, tree_combine, extract) and,
 * with reset != 0, clears it. */
int inspire_ctx_set_profiling(inspire_ctx *ctx, int enable);
/* Restricts the bracketing to the named kernels (comma-separated names as
 * above); NULL or "" brackets every kernel (the default).  Fewer events keep
 * the profiled run closer to an unprofiled one. */
int inspire_ctx_set_profiling_kernels(inspire_ctx *ctx, const char *kernels);
int inspire_ctx_profile(inspire_ctx *ctx, const char *kernel, double *total_ms, uint64_t *launches, int reset);

/* ---- kernel-variant tuning (no counterpart in the reference) ----
 *
 * Two kernels of the respond path have variants with the same results:
 *   "ntt_fwd_2048"            forward NTT at d = 2048 (database digits, keys,
 *                             the stand-alone transforms): 0 = the LDS
 *                             radix-8 kernel on the vector ALUs, 1 = a
 *                             four-step transform on the matrix cores (int8
 *                             MFMA over 7-bit limbs; moduli below 2^28);
 *   "chain_values_per_thread" the two-modulus packing chain: 4 or 8.
 * inspire_ctx_autotune times both variants of each on the device (2048
 * transforms; a 256-query InsPIRe^2 batch over a throwaway database), checks
 * that they agree word for word, keeps the faster, and returns (and, with
 * path != NULL, writes) a JSON report of the choices and times.
 * inspire_ctx_load_tuning applies a report written earlier.  Choices are
 * per context and are copied into lanes created afterwards: tune the owner
 * before creating lanes.  The report functions follow the wire encoders'
 * buffer convention (out == NULL sizes it). */
int inspire_ctx_autotune(inspire_ctx *ctx, const char *path, char *report, uint64_t capacity, uint64_t *len);
int inspire_ctx_load_tuning(inspire_ctx *ctx, const char *path);
int inspire_ctx_set_kernel_choice(inspire_ctx *ctx, const char *name, int value);
/* Measured issue rates, T operations/s over the whole device, of the
 * arithmetic the integer kernels are priced against (register-resident
 * probes, no memory traffic): tops[0] the Montgomery modular product of the
 * transforms, tops[1] its signed form, tops[2] the 64-bit multiply-add of
 * the extract kernel; with n >= 4, tops[3] the streaming read rate of HBM in
 * TB/s (16-byte loads over 1 GiB, allocated for the call).  n >= 3.  No
 * counterpart in the reference: it is the measured roof beside the
 * published one. */
int inspire_ctx_measure_rates(inspire_ctx *ctx, double *tops, uint32_t n);
int inspire_ctx_get_tuning(const inspire_ctx *ctx, char *out, uint64_t capacity, uint64_t *len);

/* ---- math entry points (NttContext, external_product) ---- */

/* NttContext::forward / inverse (src/math/ntt.rs:223-246, 308-325) over
 * `count` polynomials uint64_t[count][K][d], in place, reference NTT-domain
 * representation. */
int inspire_ntt_forward(inspire_ctx *ctx, uint64_t *polys, uint64_t count);
int inspire_ntt_inverse(inspire_ctx *ctx, uint64_t *polys, uint64_t count);
/* external_product (src/rgsw/external_product.rs:88-150) for `count`
 * independent pairs: rlwe uint64_t[count][2][K][d], rgsw
 * uint64_t[count][2*ell][2][K][d] -> out uint64_t[count][2][K][d]. */
int inspire_external_product(inspire_ctx *ctx, uint64_t count, const uint64_t *rlwe, const uint64_t *rgsw,
                             uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
