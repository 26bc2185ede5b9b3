// Respond-path drivers shared by the C ABI entry points.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

#include "context.hpp"

namespace inspire {

void hip_check(hipError_t e, const char* what);

ScaleDev scale_none();
ScaleDev scale_fwd(const inspire_ctx* ctx, const uint64_t* c);
ScaleDev scale_inv(const inspire_ctx* ctx, const uint64_t* c);
uint64_t pow2_64_mod(uint64_t q);

// u64 residues (reference layout) <-> u32 device words, through the stream.
void upload_u64(inspire_ctx* ctx, const uint64_t* src, size_t words, uint32_t* dst);
void download_u64(inspire_ctx* ctx, const uint32_t* src, size_t words, uint64_t* dst);

void db_load(inspire_ctx* ctx, uint32_t db_id, uint32_t num_shards, const uint32_t* shard_ids,
             uint32_t polys_per_shard, const uint64_t* polys);
void db_load_entries(inspire_ctx* ctx, uint32_t db_id, const uint8_t* data, uint64_t len, uint32_t entry_size);
void db_load_shard_files(inspire_ctx* ctx, uint32_t db_id, const char* dir);

void external_product_batch(inspire_ctx* ctx, uint64_t count, const uint64_t* rlwe, const uint64_t* rgsw,
                            uint64_t* out);

// How the packed variants of a call are answered.
enum class Route {
    ByVariant,  // respond_with_variant: the query's packing_mode decides (respond.rs:129-183)
    Inspiring,  // respond_inspiring / respond_seeded_inspiring (respond.rs:284-398)
    Tree,       // respond_one_packing / respond_seeded_packed (respond.rs:213-268, 428-435)
};

// Runs n queries.  variants[i] selects InsPIRe^0/1/2.
void respond_batch(inspire_ctx* ctx, uint32_t n, const inspire_query* qs, const uint32_t* variants,
                   inspire_response* rs, int32_t* status, Route route);

// RGSW rows ell..2ell-1 into rows_ws [Q][ell][2][K][N] (NTT, Montgomery
// form): queries [0, S) seeded (a expanded from seeds [S][ell][8], b in bsrc
// [S][ell][K][N]), [S, Q) full rows fsrc [Q-S][ell][2][K][N].  a_res: the
// residues of a the kernels after it read (0: all; 1: residue 0, see
// pack_rows_a_residues), except for the seeded queries t with a_full[t] != 0.
void stage_rows_from(inspire_ctx* ctx, uint32_t Q, uint32_t S, uint32_t* rows_ws, const uint32_t* seeds,
                     const uint32_t* bsrc, const uint32_t* fsrc, uint32_t* d_err, uint32_t a_res = 0,
                     const uint32_t* a_full = nullptr);

// Copies `words` device words and the 32-bit error word back into pinned
// staging (ctx->hout), waits for the stream and returns the staged words.
const uint32_t* fetch_results(inspire_ctx* ctx, const uint32_t* d_out, size_t words, const uint32_t* d_err,
                              uint32_t* err);

// Device address of the digits of the shard in `slot`.
uint64_t shard_digits(const Database& db, uint32_t slot);

// InsPIRe^0 on staged NTT-domain rows: out [Q][1 + cols][2][K][N].
void run_respond0(inspire_ctx* ctx, uint32_t cols, uint32_t Q, const uint32_t* rows, const uint64_t* d_qdig,
                  uint32_t* out);

// Scratch layout of a device-resident batch in one buffer (byte offsets,
// each part 256-byte aligned): digit addresses and error word, transformed
// rows, and for InspiRING the packing scratch (PackWs) and transformed keys.
// tables is null when the InspiRING tables of the column count are not built
// yet (then the packing parts are not sized).
struct PackTables;
struct DeviceBatchPlan {
    bool tree = false, packed = false;
    uint32_t cols = 0;
    const PackTables* tables = nullptr;
    size_t o_qdig = 0, o_rows = 0, o_rpre = 0, o_bpart = 0, o_cols = 0, o_y = 0, bytes = 0;
};
// Checks the batch as the entry point does (throws its errors) and lays out its scratch.
DeviceBatchPlan device_batch_plan(inspire_ctx* ctx, const inspire_device_batch& b);

// Same work for a batch whose inputs are resident in device memory.  With
// scratch == nullptr the device scratch comes from the context's workspaces
// and the host staging from its pinned buffer; otherwise scratch holds
// device_batch_plan(ctx, b).bytes and host_meta count + 2 words of host
// memory, and nothing is allocated.
void respond_batch_device(inspire_ctx* ctx, const inspire_device_batch& b, uint8_t* scratch = nullptr,
                          uint64_t* host_meta = nullptr);

}  // namespace inspire
