// Match-distance capture for anonymized statistics: the code/mask dot-product
// shares of opened matches (and of requested non-matching pairs) with their
// match id, as openResults / storeAnonStatsWithIndexMapping write them
// (dot/kernel.cu:54-116,153-210; buffers of server/anon_stats.rs:13-123).
#pragma once
#include "common.hpp"

namespace irismpc {

struct AnonCapture {
  uint16_t* code_a = nullptr;  // capacity entries each
  uint16_t* code_b = nullptr;
  uint16_t* mask_a = nullptr;
  uint16_t* mask_b = nullptr;
  uint64_t* idx = nullptr;     // match id
  uint32_t* counter = nullptr; // entries seen (may exceed capacity)
  size_t capacity = 0;
  uint64_t batch_id = 0;
  size_t max_query_length = 0;  // match id = batch_id * max_db * max_query + db * max_query + row
  size_t max_db_length = 0;
  const uint64_t* targets = nullptr;  // db * (max_query_length / 31) + query: captured
  size_t n_targets = 0;               // even when not matching (reauth targets)
  __host__ __device__ bool enabled() const { return capacity > 0 && counter != nullptr; }
};

// The chunk's dot-product shares (replicated a = own, b = previous party's),
// element q * chunk_len + column, as the circuits consumed them.
struct DotViews {
  const uint16_t* code_a = nullptr;
  const uint16_t* code_b = nullptr;
  const uint16_t* mask_a = nullptr;
  const uint16_t* mask_b = nullptr;
};

__device__ inline void anon_store(const AnonCapture& c, const DotViews& d, size_t e, size_t db,
                                  size_t row) {
  const uint32_t k = atomicAdd(c.counter, 1u);
  if (k >= c.capacity) return;
  c.idx[k] = c.batch_id * (uint64_t)(c.max_db_length * c.max_query_length) +
             (uint64_t)db * c.max_query_length + row;
  c.code_a[k] = d.code_a[e];
  c.code_b[k] = d.code_b[e];
  c.mask_a[k] = d.mask_a[e];
  c.mask_b[k] = d.mask_b[e];
}

// Non-matching target pairs of one opened chunk: thread (target, rotation).
// Chunk column p holds DB row offset + p (index_map == nullptr) or
// index_map[p]; elements whose opened bit is set were captured as matches.
struct AnonTargetArgs {
  AnonCapture cap;
  DotViews dots;
  const uint64_t* a;
  const uint64_t* b;
  const uint64_t* c;
  size_t chunk_len, real_chunk_len, offset, batch;
  const uint32_t* index_map;
};
void launch_anon_targets(const AnonTargetArgs& a, hipStream_t s);

}  // namespace irismpc
