// One party's full-DB scan of a query batch: the chunk loop of
// ServerActor::compare_query_against_db_and_self_prefiltering
// (iris-mpc-gpu/src/server/actor.rs:2246-2589), with the DB resident in HBM so
// the prefetch stage disappears.
#pragma once
#include "engine.hpp"

namespace irismpc {

struct PartyScan {
  ShareDB* codes;
  ShareDB* masks;
  Circuits* circuits;
  DistanceComparator* comparator;
  const DbSlice* code_db;
  const DbSlice* mask_db;
  const QuerySet* code_query;
  const QuerySet* mask_query;
};

// For every chunk of kDbChunk rows: code/mask dot products with the fused
// correction and zero-share masking, reshare, compare against threshold_a,
// open and OR the opened bits into `bitmap` (bit total_db*(q/31) + db); with
// a distance capture set on the comparator, matches' distances are captured.
// Collective: the three parties call it concurrently (one thread each).
void scan_db(const PartyScan& p, size_t db_size, size_t batch_size, uint32_t threshold_a,
             uint64_t* bitmap, hipStream_t s);

// One chunk of that loop: DB rows [offset, offset + min(kDbChunk, db_size -
// offset)) against the whole query set (batch_size queries, the query sets'
// rows; the last chunk is padded to 64 rows, or 512 when the rows are not a
// multiple of 8). slot_of (device, optional) maps batch query j to bitmap row
// slot_of[j]; bitmap rows are row_bits bits apart (0: db_size, the layout of
// scan_db). Collective like scan_db.
void scan_chunk(const PartyScan& p, size_t offset, size_t db_size, size_t batch_size,
                uint32_t threshold_a, uint64_t* bitmap, const uint32_t* slot_of, size_t row_bits,
                hipStream_t s);
void scan_chunk_trio(const PartyScan p[3], size_t offset, size_t db_size, size_t batch_size,
                     uint32_t threshold_a, uint64_t* const bitmaps[3],
                     const uint32_t* const slot_of[3], size_t row_bits, hipStream_t const s[3]);
// n_chunks consecutive chunks from `offset` in one step (their results side by
// side in the result buffer: n_chunks x 32,768 x rows must fit it).
constexpr size_t kMaxStepChunks = 16;
void scan_chunks(const PartyScan& p, size_t offset, size_t n_chunks, size_t db_size,
                 size_t batch_size, uint32_t threshold_a, uint64_t* bitmap,
                 const uint32_t* slot_of, size_t row_bits, hipStream_t s);
void scan_chunks_trio(const PartyScan p[3], size_t offset, size_t n_chunks, size_t db_size,
                      size_t batch_size, uint32_t threshold_a, uint64_t* const bitmaps[3],
                      const uint32_t* const slot_of[3], size_t row_bits, hipStream_t const s[3]);

// The DB half of compare_query_against_db_subset_and_self (actor.rs:1900-2244):
// the subset rows ids[0..n_ids) (device u32, DB row ids below total_db_len)
// are gathered chunk by chunk (DB_CHUNK_SIZE rows) from the HBM DB into the
// subset slices (prefetch_subset), compared against the whole query set at
// threshold_a, and the opened bits set at the rows' DB ids
// (openResultsWithIndexMapping), with the rotation capture. n_ids = 0 does
// nothing. Collective like scan_db.
// anon_threshold != 0 (and a distance capture set on the comparator): each
// chunk is first compared at anon_threshold and only its distances captured
// (the db_threshold_anon_stats pass, actor.rs:2040-2120).
void scan_subset(const PartyScan& p, DbSlice& code_sub, DbSlice& mask_sub, const uint32_t* ids,
                 size_t n_ids, size_t total_db_len, size_t batch_size, uint32_t threshold_a,
                 uint64_t* bitmap, hipStream_t s, uint32_t anon_threshold = 0);
void scan_subset_trio(const PartyScan p[3], DbSlice* const code_sub[3], DbSlice* const mask_sub[3],
                      const uint32_t* const ids[3], size_t n_ids, size_t total_db_len,
                      size_t batch_size, uint32_t threshold_a, uint64_t* const bitmaps[3],
                      hipStream_t const s[3], uint32_t anon_threshold = 0);

// compare_query_against_self (actor.rs:1789-1897): the batch's query rows
// (all rotations) against the batch's own irises in DB form (the query sets'
// insert slices) at threshold_a. Bit row_bits * q + j (row_bits >= batch;
// 0 = batch) = some rotation of query q matches query j. The reference's batch
// DB holds all 31 rotations of each query, of which mergeBatchResults reads
// only the unrotated one; only those columns are computed here.
void scan_batch_self(const PartyScan& p, size_t batch_size, uint32_t threshold_a,
                     uint64_t* bitmap, size_t row_bits, hipStream_t s);
void scan_batch_self_trio(const PartyScan p[3], size_t batch_size, uint32_t threshold_a,
                          uint64_t* const bitmaps[3], size_t row_bits, hipStream_t const s[3]);

// Runs the three parties' scans from one call: from this thread, with the six
// share GEMMs of a chunk in order on s[0] and the circuits on the exchange's
// protocol stream (one host thread per party when a distance capture is set).
void scan_db_trio(const PartyScan p[3], size_t db_size, size_t batch_size, uint32_t threshold_a,
                  uint64_t* const bitmaps[3], hipStream_t const s[3]);

}  // namespace irismpc
