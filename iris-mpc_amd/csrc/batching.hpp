// Batching form: consecutive query batches of uneven size (1-256 queries x 31
// rotations) packed into one query set per pack, so one share-GEMM launch per
// DB chunk serves all of a pack's batches. Each batch keeps the output layout
// and ChaCha12 mask counters of its own plain run (the segment table of the
// packed launch), so both forms give identical shares. The reference runs one
// batch at a time, its query set padded to max_batch_size x 31 rows
// (iris-mpc-gpu/src/server/actor.rs:895-901); packing is this library's.
#pragma once
#include <memory>
#include <vector>

#include "engine.hpp"

namespace irismpc {

// Greedy packing of consecutive batches: a pack holds at most max_queries
// queries (a larger batch gets a pack of its own) and at most max_batches
// batches. Returns the packs as lists of batch indices.
std::vector<std::vector<size_t>> plan_packs(const uint32_t* sizes, size_t n, size_t max_queries,
                                            size_t max_batches);
// ChaCha blocks a whole scan of n_rows DB rows uses on a remote engine, in
// ranges of `chunk` rows with the last padded to 64 rows (as the runners
// launch it): mask + reshare OTP per chunk.
uint64_t scan_counter_advance(size_t n_rows, size_t chunk, size_t rows);
// The serving form's step size from measured seconds per chunk (k[i],
// seconds[i]): steps longer than max_step_s are left out (a request arriving
// meanwhile waits up to a whole step to join the active set; when every step
// is longer, the smallest k); of the rest, the smallest k within `tol` of the
// fastest per chunk.
uint32_t pick_step_chunks(const uint32_t* k, const double* seconds_per_chunk, size_t n, double tol,
                          double max_step_s);

// Plain and packed runs of the same batches over one matrix (code or mask) of
// one party. outs[i] (device, caller-owned) receives batch i's results as its
// own dot_reduce_range over the DB lays them out: the last chunk's results
// stay (chunk x rows_i u16, + 64).
class BatchRunner {
 public:
  BatchRunner(ShareDB* engine, const DbSlice* db, size_t n_rows, uint16_t multiplier,
              const uint32_t* sizes, size_t n_batches, const uint16_t* shares,
              uint16_t* const* outs, size_t chunk, size_t max_queries, size_t max_batches,
              hipStream_t s);
  // New pack size: rebuilds the plan and the packed query sets (the old sets
  // are freed first; synchronises the device).
  void replan(size_t max_queries, hipStream_t s);
  const std::vector<std::vector<size_t>>& plan() const { return plan_; }
  size_t max_queries() const { return max_queries_; }
  // The packed schedule over the first n_rows DB rows (0: all): every pack,
  // every chunk, one launch each, with the counters each batch's plain run
  // starting at ctr0 would use. No allocation or synchronisation, so it can
  // be captured into a graph.
  void issue_packed(uint64_t ctr0, size_t n_rows, hipStream_t s);
  // One batch's plain run over the whole DB at the engine's counter (which
  // it advances), chunk by chunk.
  void issue_plain(size_t batch, hipStream_t s);
  // Persistent form of the packed schedule: one launch per chunk covering all
  // packs (share_gemm_rot_multi_kernel, a block per CU looping over the
  // chunk's tiles), with the outputs and counters of issue_packed(ctr0).
  // prepare_persistent builds the launch descriptors in device memory
  // (synchronises); issue_persistent allocates and synchronises nothing, so
  // it can be captured into a graph. All packs run the 4-query tile.
  void prepare_persistent(uint64_t ctr0, hipStream_t s);
  void issue_persistent(hipStream_t s);

 private:
  ShareDB* e_;
  const DbSlice* db_;
  size_t n_rows_, chunk_, max_queries_, max_batches_;
  uint16_t mult_;
  std::vector<uint32_t> sizes_;
  std::vector<size_t> first_;  // first query of batch i in `shares_`
  std::vector<uint16_t> shares_;
  std::vector<uint16_t*> outs_;
  std::vector<std::vector<size_t>> plan_;
  std::vector<std::unique_ptr<QuerySet>> q_plain_, q_packed_;
  // persistent form: descriptors [chunk][pack] and running tile counts
  DevBuf<GemmArgs> descs_;
  DevBuf<uint32_t> tile_end_;
  size_t n_ranges_ = 0;
  int grid_ = 0;
};

}  // namespace irismpc
