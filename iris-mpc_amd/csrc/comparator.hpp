// DistanceComparator device helpers (merge of eye bitmaps, buffer fills).
#pragma once
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
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s);

}  // namespace irismpc
