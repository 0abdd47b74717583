// DistanceComparator device helpers (merge of eye bitmaps, buffer fills).
#pragma once
#include "anon.hpp"
#include "common.hpp"

namespace irismpc {

struct MergeArgs {
  const uint64_t* left;
  const uint64_t* right;
  const uint64_t* or_policy;  // nullptr = AND policy
  uint32_t* final_results;
  uint32_t* match_counters;
  uint32_t* all_matches;
  uint32_t* mcl;
  uint32_t* mcr;
  uint32_t* prl;
  uint32_t* prr;
  size_t query_len;   // queries (batch), not rotations
  size_t db_len;
  size_t max_db_len;  // row stride of the OR-policy bitmap
};

void launch_merge(const MergeArgs& a, hipStream_t s);

struct BatchMergeArgs {
  const uint64_t* left;
  const uint64_t* right;
  uint32_t* final_results;
  uint32_t* match_counters;
  uint32_t* all_matches;
  size_t batch;
  size_t row_bits;
};
void launch_merge_batch(const BatchMergeArgs& a, hipStream_t s);

// Candidate DB ids of a match bitmap (partialDbResults, dot/kernel.cu:212-239,
// and the host-side unique + sort of get_partial_results,
// distance_comparator.rs:545-648): every id d < db_size with bit
// row_bits * q + d set for some query q < batch, plus the extra ids (OR-rule
// and reauth targets, actor.rs:946-970) below db_size, as one ascending list
// without repeats. Scratch: cand (ceil(db_size / 64) words) and tiles
// (ceil(db_size / 16384) + 1 words); *count receives the list length.
struct CandidateArgs {
  const uint64_t* bitmap;
  size_t row_bits;
  size_t batch;
  size_t db_size;
  const uint32_t* extra;  // device, n_extra ids (any order, repeats allowed)
  size_t n_extra;
  uint64_t* cand;
  uint32_t* tiles;
  uint32_t* ids;
  uint32_t* count;
};
void launch_candidates(const CandidateArgs& a, hipStream_t s);
// The candidate list of launch_candidates with its scratch in a file-scope
// device array (no allocation, no synchronisation): extra (device, n_extra)
// ids join the bitmap's; ids (device) receive the list, *count (device) its
// length. db_size <= kNoAllocMaxDb.
constexpr size_t kNoAllocMaxDb = size_t(1) << 27;
constexpr int kMaxDevices = 64;
void partial_results_noalloc(const uint64_t* bitmap, size_t db_size, size_t batch, size_t row_bits,
                             const uint32_t* extra, size_t n_extra, uint32_t* ids, uint32_t* count,
                             hipStream_t s);

// openResultsWithIndexMapping (dot/kernel.cu:118-151): opened bit (q, j) of a
// subset chunk sets bitmap bit total_db_len * slot + index_map[j] and records
// (q / 31, index_map[j], q % 31 - 15) in the rotation capture (first
// capacity matches; *counter counts them all).
struct MappedOpenArgs {
  const uint64_t* a;
  const uint64_t* b;
  const uint64_t* c;
  uint64_t* bitmap;
  size_t n_words;
  size_t chunk_len;
  size_t query_len;  // real batch * 31
  size_t real_chunk_len;
  size_t total_db_len;
  const uint32_t* index_map;
  const uint32_t* slot_of;  // optional
  uint32_t* counter;
  uint32_t* cap_query;
  uint32_t* cap_db;
  int8_t* cap_rot;
  size_t capacity;
  int set_bits;      // 0: capture distances only (the anonymized-statistics pass)
  AnonCapture anon;  // optional match-distance capture
  DotViews dots;
};
void launch_open_mapped(const MappedOpenArgs& a, hipStream_t s);
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s);

}  // namespace irismpc
