// Request loops over the three parties' HBM-resident DB shares: the serving
// form (continuous batching at chunk-step granularity) and the reference's
// per-batch form (ServerActor's batch after batch, every batch padded to
// max_batch x 31 query rows, iris-mpc-gpu/src/server/actor.rs:584-644,895-901)
// on the same engines, so the two can be compared on the same requests.
//
// Requests (one iris each, raw shares of the three parties):
// - uniqueness: one query x 31 rotations against the whole DB at threshold_a;
//   the result is the ascending list of matching DB rows;
// - reauth: one query x 31 rotations against one target row (a 64-row window
//   on its own engine set and streams, concurrent with the scan); the result
//   is 0 / 1;
// - insertion: the iris appended to the DB (result: its row);
// - deletion: the row overwritten with shares of the default iris (code 0,
//   mask all ones), as get_dummy_shares_for_deletion does
//   (iris-mpc-common/src/iris_db/mod.rs:10-23); result: the row.
//
// Serving form: the scan of actor.rs:2246-2589 becomes a ring of chunk steps.
// Uniqueness requests join the active query set between steps while a slot is
// free and free HBM stays above a reserve plus the request's footprint (with
// nothing active one request always joins); a request leaves once it has
// covered every chunk of the DB. After a write (insertion or deletion)
// between steps, the written chunk is scanned again at once for the whole
// active set (its stale bits cleared first; an insertion into a new chunk
// adds that chunk), so each answer is that of the DB as it stands at the
// request's last step: the DB after exactly Result::db_version of the loop's
// writes, in the order the loop applied them (as a batch of the reference
// sees the writes before it). Requests in flight together are compared with
// each other as the reference compares a batch with itself
// (compare_query_against_self and mergeBatchResults, actor.rs:1789-1897,
// kernel.cu:352-386): after a refill that admitted uniqueness requests, every
// query of the active set against the set's own irises, and a request reports
// the earlier-admitted requests still scanning at its admission that it
// matches (Result::batch_matches). The per-batch form does the same for each
// batch. Every step runs,
// for all active queries, one share GEMM per matrix and party over 1-16
// chunks, the reshare, the threshold circuits and the open into each
// request's bitmap row (slot map). Nothing between steps drains the queued
// steps: query refills, slot maps, slot clearing, bitmap readouts and DB
// writes are ordered on the parties' scan streams.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "actor.hpp"

namespace irismpc {

enum ServingKind : int { kReqUniqueness = 0, kReqReauth = 1, kReqInsert = 2, kReqDelete = 3 };
enum ServingPolicy : int { kPolicyContinuous = 0, kPolicyPerBatch = 1 };
enum ServingState : int { kReqPending = 0, kReqDone = 1, kReqFailed = -1 };

// Page-locked host memory (hipHostMalloc).
template <typename T>
struct PinnedBuf {
  T* ptr = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  explicit PinnedBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    release();
    IRIS_HIP_CHECK(hipHostMalloc((void**)&ptr, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  ~PinnedBuf() { release(); }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
};

// The engines of the three parties for one query-set size: code and mask
// ShareDB, Circuits and DistanceComparator per party over one exchange, a
// stream per party, and a persistent code / mask query set per party that
// requests refill in place. Engine seeds are derive_seed(party seed,
// nonce_base + 0 / 1 / 5) for code, mask and circuits: every engine set takes
// its own nonce_base, so no two draw the same keystream (the reference takes
// a fresh kdf nonce per engine, actor.rs:382-388).
struct EngineTrio {
  // chunk_rows: DB rows per scan step (the result and circuit buffers' size)
  EngineTrio(int device, size_t max_queries, size_t db_capacity, const ChaChaKey own[3],
             const ChaChaKey prev[3], int tag0, uint64_t nonce_base, size_t chunk_rows = kDbChunk);
  ~EngineTrio();
  EngineTrio(const EngineTrio&) = delete;
  EngineTrio& operator=(const EngineTrio&) = delete;
  void sync() const;
  // the parties' scan arguments over the given DB slices
  void parties(DbSlice* const code_dbs[3], DbSlice* const mask_dbs[3], PartyScan out[3]) const;

  size_t max_queries;
  std::unique_ptr<Exchange> ex;
  std::unique_ptr<ShareDB> codes[3], masks[3];
  std::unique_ptr<Circuits> circ[3];
  std::unique_ptr<DistanceComparator> cmp[3];
  std::unique_ptr<QuerySet> qc[3], qm[3];
  hipStream_t s[3] = {nullptr, nullptr, nullptr};
};

struct ServingStats {
  uint64_t steps = 0;            // chunk steps run (serving and per-batch forms)
  uint64_t max_concurrent = 0;   // largest active set (serving form)
  uint64_t admission_waits = 0;  // admissions deferred by the free-HBM reserve
  uint64_t batches = 0;          // batches run (per-batch form)
  uint64_t refills = 0;          // query-set refills (active-set changes)
  // host seconds of the loop thread: refills (staging and queueing), waits
  // for the queued steps (at most two stay queued), queueing the steps,
  // free-HBM checks, bitmap decoding
  double t_refill = 0, t_wait = 0, t_enqueue = 0, t_admit = 0, t_decode = 0;
};

class Serving {
 public:
  // party_seeds[p]: party p's seed; its engines' prev seed is party p - 1's.
  // The DB slices (allocated by any ShareDB of the matching code length) hold
  // n_rows rows and room for `capacity`; they stay owned by the caller.
  Serving(int device, const ChaChaKey party_seeds[3], DbSlice* const code_dbs[3],
          DbSlice* const mask_dbs[3], size_t n_rows, size_t capacity, size_t max_active,
          uint32_t threshold_a);
  ~Serving();
  Serving(const Serving&) = delete;
  Serving& operator=(const Serving&) = delete;

  // Starts the request loop (and the reauth worker); requests submitted
  // before start wait for it. max_batch: the per-batch form's padding
  // (1 .. max_active); reserve_bytes: free HBM kept by admission.
  void start(int policy, size_t max_batch, size_t reserve_bytes);
  // Waits until every submitted request has finished, then stops the loop. A
  // failure of the loop itself is raised here. After timeout_s (<= 0: none)
  // the loop is stopped with the unfinished requests failed, and the object
  // serves no further requests.
  void finish(double timeout_s);
  bool running() const { return running_; }

  // code: [3][12800] u16, mask: [3][6400] u16 (the parties' raw shares;
  // unused for deletions); arrival: seconds on now()'s clock (< 0: now). A
  // request with a later arrival is held until then, so a caller can submit
  // a whole arrival schedule at once (requests of one arrival time are then
  // seen together, as a burst is).
  uint64_t submit(int kind, const uint16_t* code, const uint16_t* mask, int64_t target,
                  double arrival);
  struct Result {
    int state = kReqPending;
    int64_t value = -1;  // reauth 0/1, insertion / deletion row, uniqueness: match count
    std::vector<uint32_t> matches;
    double arrival = 0, done = -1;
    size_t db_rows = 0;  // DB rows the request was checked against
    // DB writes applied before the answer (a write: its own position, from 1):
    // the answer is that of the DB after the first db_version writes
    uint64_t db_version = 0;
    // uniqueness: the requests admitted before it and scanning with it (the
    // same batch in the per-batch form) whose iris it matches, by id
    std::vector<uint64_t> batch_matches;
    std::string error;
  };
  // false when the id is unknown
  bool result(uint64_t id, Result* out) const;
  // Forgets a finished request (its result can no longer be read); false
  // while it is pending. A serving process releases every request it has
  // read, so the records do not grow without bound.
  bool release(uint64_t id);
  double now() const;
  ServingStats stats() const;
  size_t n_rows() const;
  size_t max_active() const { return max_active_; }

  // Plain per-request form (loop stopped): one uniqueness query alone, a full
  // scan in steps of chunks_per_step(1, ...); returns the matching rows.
  std::vector<uint32_t> run_alone(const uint16_t* code, const uint16_t* mask);
  // Seconds of one scan step of `batch` default-iris queries over n_chunks
  // chunks from the ring's start, best of `reps` after one warm-up (loop
  // stopped; the slots it writes are cleared when requests take them).
  double time_step(size_t batch, size_t n_chunks, int reps);
  // Admission of the serving form: requests join while fewer than
  // active_cap are active, and all waiting ones join (up to the slots) while
  // more than queue_threshold wait. The default (active_cap = max_active)
  // admits every request a slot allows. A small cap keeps each active
  // request's steps short (its latency is the DB's chunks times the step
  // time of the active set); the threshold trades that for throughput when a
  // queue builds up.
  void set_admission(size_t active_cap, size_t queue_threshold);
  // Step sizes: table[i] chunks per step for active counts from buckets[i]
  // up to the next bucket (n = 0 clears the table: a fixed rule applies).
  void set_step_table(const uint32_t* buckets, const uint32_t* chunks, size_t n);
  size_t chunks_per_step(size_t batch, size_t chunk_index, size_t n_chunks) const;
  static size_t n_chunks_of(size_t rows) { return std::max<size_t>(1, ceil_div(rows, kDbChunk)); }

 private:
  struct Req {
    uint64_t id = 0;
    int kind = 0;
    std::vector<uint16_t> code, mask;
    int64_t target = -1;
    Result res;
    // loop bookkeeping
    int slot = -1;
    size_t need = 0, n_covered = 0;
    std::vector<uint8_t> covered;
    hipEvent_t readout = nullptr;
    uint64_t self_round = 0;  // the self-comparison holding its row (0: none)
  };
  struct StepEvents {
    hipEvent_t e[3];
  };

  void loop_main(int policy, size_t max_batch, size_t reserve_bytes);
  void loop_continuous(size_t reserve_bytes);
  void loop_per_batch(size_t max_batch);
  void reauth_main();
  void serve_reauth(Req* r);
  void apply_write(Req* r);
  // one step over a chunk written since the last step, for every active
  // request (the chunk's stale bits cleared first), which then has it covered
  void patch_step(const std::vector<Req*>& active, size_t chunk, size_t rows);
  void clear_chunk(int slot, size_t chunk);
  void complete(Req* r, int state, const std::string& err = std::string());
  bool take_incoming(std::vector<Req*>* out, bool wait);
  void fail_all(const std::string& msg);

  // device steps (loop thread, or the caller while stopped)
  // returns the staging ring entry that holds the set's raw shares
  int refill_scan(const std::vector<const Req*>& reqs, size_t batch, size_t rows);
  // The self-comparison after a refill from ring entry k: ids[j] is the
  // uniqueness request at position j of the set (0 elsewhere); the queries
  // at positions [q0, ids.size()) (31 rotations each) against the set's
  // irises, on the comparison engines and streams (beside the scan); each
  // fresh request at position b gets the ids[j], j < b, whose iris it
  // matches.
  void self_compare(int k, const std::vector<uint64_t>& ids, size_t q0,
                    const std::vector<std::pair<Req*, size_t>>& fresh);
  // decodes the finished self-comparisons (all up to round `upto`, waiting)
  void collect_self(uint64_t upto);
  void set_slots(const std::vector<int>& slots);
  void clear_slot(int slot);
  void step(size_t batch, size_t chunk_index, size_t n_chunks, size_t db_rows);
  void drain_steps();
  hipEvent_t begin_readout(int slot);
  std::vector<uint32_t> readout_rows(int slot, size_t db_rows) const;
  hipEvent_t event();
  void recycle(hipEvent_t e);

  int device_;
  size_t capacity_, max_active_, words_;
  uint32_t threshold_a_;
  DbSlice* code_dbs_[3];
  DbSlice* mask_dbs_[3];
  std::unique_ptr<EngineTrio> scan_, small_;
  PartyScan scan_p_[3], small_p_[3];
  DevBuf<uint64_t> bitmaps_[3], small_bitmaps_[3];
  DevBuf<uint32_t> slot_maps_[3];
  // default-iris shares: device rows (deletions) and host copies (padding)
  DevBuf<uint16_t> dummy_code_[3], dummy_mask_[3];
  std::vector<uint16_t> dummy_code_host_, dummy_mask_host_;
  // staging rings in pinned memory, each entry reused once its events passed
  static constexpr int kRefillRing = 3, kSlotRing = 4, kWriteRing = 16;
  PinnedBuf<uint16_t> refill_code_[kRefillRing], refill_mask_[kRefillRing];
  std::vector<hipEvent_t> refill_ev_[kRefillRing];
  int refill_next_ = 0;
  PinnedBuf<uint32_t> slot_host_[kSlotRing];
  std::vector<hipEvent_t> slot_ev_[kSlotRing];
  int slot_next_ = 0;
  std::vector<int> slots_;  // what the device slot maps hold
  PinnedBuf<uint16_t> write_code_[kWriteRing], write_mask_[kWriteRing];
  DevBuf<uint16_t> write_dev_code_[kWriteRing], write_dev_mask_[kWriteRing];
  std::vector<hipEvent_t> write_ev_[kWriteRing];
  int write_next_ = 0;
  hipEvent_t last_write_[3] = {nullptr, nullptr, nullptr};  // DB writes, per scan stream
  PinnedBuf<uint64_t> readout_;  // party 0's bitmap row per slot
  PinnedBuf<uint16_t> small_code_, small_mask_;
  struct SelfRound {
    uint64_t seq = 0;
    int buf = 0;
    hipEvent_t done = nullptr;
    std::vector<uint64_t> ids;
    size_t q0 = 0;
    std::vector<std::pair<Req*, size_t>> fresh;
  };
  static constexpr int kSelfRing = 4;
  static constexpr size_t kSelfChunk = 512;  // DB rows of the comparison engines' step
  std::unique_ptr<EngineTrio> self_;         // the comparison engines (own keystreams and streams)
  std::unique_ptr<DbSlice> self_code_db_[3], self_mask_db_[3];  // the set's irises, one row each
  PartyScan self_p_[3];
  DevBuf<uint16_t> self_raw_code_[3], self_raw_mask_[3];  // staging of the set's raw shares
  DevBuf<uint64_t> self_bitmaps_[3];     // one 64-bit row per compared query
  PinnedBuf<uint64_t> self_host_[kSelfRing];
  std::deque<SelfRound> self_rounds_;
  uint64_t self_seq_ = 0;
  std::deque<StepEvents> inflight_;
  std::vector<hipEvent_t> event_pool_;
  std::vector<hipEvent_t> recycle_later_;  // events of released requests (destroyed at the end)
  std::vector<uint32_t> step_buckets_, step_chunks_;
  size_t active_cap_ = 64, queue_threshold_ = 0;

  // requests and threads
  std::chrono::steady_clock::time_point t0_;
  mutable std::mutex mu_;        // requests, queues, stats
  std::condition_variable cv_;   // new submissions / stop
  std::condition_variable done_cv_;
  std::condition_variable reauth_cv_;
  std::mutex db_mu_;             // DB writes vs the reauth worker's reads
  std::map<uint64_t, std::unique_ptr<Req>> reqs_;
  std::vector<Req*> incoming_;
  std::deque<Req*> reauth_q_;
  uint64_t next_id_ = 1, n_submitted_ = 0, n_finished_ = 0;
  uint64_t n_writes_ = 0;  // DB writes applied (the DB's version)
  size_t n_rows_;
  bool running_ = false, draining_ = false, reauth_stop_ = false;
  std::atomic<bool> abort_{false};  // a timed-out finish: leave at once
  std::exception_ptr loop_error_;
  std::thread loop_, reauth_;
  ServingStats stats_;
};

}  // namespace irismpc
