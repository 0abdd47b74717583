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
// open and OR the opened bits into `bitmap` (bit total_db*(q/31) + db).
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

// Runs the three parties' scans from one call (one host thread per party).
void scan_db_trio(const PartyScan p[3], size_t db_size, size_t batch_size, uint32_t threshold_a,
                  uint64_t* const bitmaps[3], hipStream_t const s[3]);

}  // namespace irismpc
