// Host-side engines mirroring the reference GPU server's matching path:
// ShareDB (iris-mpc-gpu/src/dot/share_db.rs), Circuits
// (threshold_ring/protocol.rs), DistanceComparator (dot/distance_comparator.rs)
// and the NCCL communicator replaced by an in-process three-party Exchange.
#pragma once
#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "circuits.hpp"
#include "common.hpp"
#include "share_gemm.hpp"

namespace irismpc {

// RAII device allocation.
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count, bool zero = true) { alloc(count, zero); }
  void alloc(size_t count, bool zero = true) {
    release();
    n = count;
    if (count) {
      IRIS_HIP_CHECK(hipMalloc(&ptr, count * sizeof(T) + 256));
      if (zero) {
        // hipMemset runs on the null stream, which does not order with the
        // engines' non-blocking streams: finish it before any of them can
        // write the buffer (a lazily allocated result slot is first written
        // by a GEMM on such a stream right after this call)
        IRIS_HIP_CHECK(hipMemset(ptr, 0, count * sizeof(T) + 256));
        IRIS_HIP_CHECK(hipStreamSynchronize(nullptr));
      }
    }
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

struct Event {
  hipEvent_t e = nullptr;
  Event() { IRIS_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
};

// Opt-in timing of an engine's kernel launches: begin/end record HIP events
// on the launch stream around one launch (nothing waits at record time);
// collect() waits for the recorded events and returns the number of timed
// launches and their summed durations since the previous collect.
class LaunchTimer {
 public:
  LaunchTimer() = default;
  ~LaunchTimer();
  LaunchTimer(const LaunchTimer&) = delete;
  LaunchTimer& operator=(const LaunchTimer&) = delete;
  void enable(bool on) { on_ = on; }
  bool on() const { return on_; }
  void begin(hipStream_t s);
  void end(hipStream_t s);
  size_t collect(double* total_ms);

 private:
  struct Pair {
    hipEvent_t a = nullptr, b = nullptr;
  };
  std::vector<Pair> ev_;
  size_t used_ = 0;
  bool on_ = false;
};

// The link between the three parties on one GPU (replaces NcclComm,
// helpers/comm.rs). Party p sends to next = p+1 and receives from prev = p-1.
// Operations that the reference performs as matched send/recv pairs are
// collectives here: every party calls with its own arguments and stream; the
// last arrival performs the device-to-device exchange for all three.
class Exchange {
 public:
  explicit Exchange(int device, double timeout_s = 300.0);
  ~Exchange();
  int device() const { return device_; }
  hipStream_t protocol_stream() const { return protocol_; }
  // Rendezvous: blocks until all three parties called with the same op id.
  void collective(int party, uint64_t op, void* arg, const std::function<void(void* const*)>& fn);
  // Orders `protocol` after the three party streams and the party streams
  // after `protocol` (the exchange barrier of a joint launch).
  void join_streams(hipStream_t const s[3]);
  void fork_streams(hipStream_t const s[3]);

 private:
  int device_;
  double timeout_s_;
  hipStream_t protocol_ = nullptr;
  Event ev_party_[3], ev_proto_;
  std::mutex mu_;
  std::condition_variable cv_;
  void* args_[3] = {nullptr, nullptr, nullptr};
  uint64_t ops_[3] = {0, 0, 0};
  int arrived_ = 0;
  uint64_t generation_ = 0;
  std::exception_ptr error_;
  uint64_t error_gen_ = ~0ull;
};

// SlicedProcessedDatabase analogue: one matrix (code or mask) of one party,
// resident in HBM in MFMA fragment order, plus per-row limb sums.
class ShareDB;
struct DbSlice {
  int code_length = 0;
  size_t capacity = 0;   // rows addressable (padded)
  DevBuf<int8_t> limbs;  // tiled i8 limbs
  DevBuf<uint32_t> sum0, sum1;
  DevBuf<uint16_t> staging;  // host upload staging
};

// CudaVec2DSlicerU8 + sums analogue for a query batch.
struct QuerySet {
  int code_length = 0;
  int rows = 0;      // query_length rows
  DevBuf<int8_t> limbs;
  DevBuf<uint32_t> sum0, sum1;
  // padded unrotated limb planes; present when the rows are the 31 rotations
  // of each query (built from raw shares on the device)
  DevBuf<int8_t> rot;
  // The batch's queries in DB form (their raw shares, unrotated, not
  // interpolated: the centre rotation of the reference's code_query_insert),
  // present when uploaded from raw shares. Source of write_at_index and the
  // DB side of the batch self-comparison.
  int batch = 0;
  std::unique_ptr<DbSlice> insert;
  // what the buffers hold (a refill with fewer rows or queries reuses them)
  size_t cap_rows = 0, cap_batch = 0;
};

class ShareDB {
 public:
  ShareDB(int device, int peer_id, size_t max_db_length, size_t query_length, int code_length,
          const ChaChaKey& seed_own, const ChaChaKey& seed_prev, Exchange* ex, int tag);
  ~ShareDB();
  std::unique_ptr<DbSlice> alloc_db(size_t max_db_length) const;
  // load_single_record_from_db for n consecutive rows (host u16 -> HBM)
  void load_records_u16(DbSlice& db, size_t index, size_t n, const uint16_t* host, hipStream_t s) const;
  // load_single_record_from_s3 (host even/odd limbs -> HBM)
  void load_records_limbs(DbSlice& db, size_t index, size_t n, const uint8_t* l0, const uint8_t* l1,
                          hipStream_t s) const;
  void write_at_index(DbSlice& db, const QuerySet& q, size_t src_index, size_t dst_index,
                      hipStream_t s) const;
  // stored records already on the device (layout above): the field at
  // (off_lo, off_hi) of each record with serial id in [1, max_serial] -> row id - 1
  void load_records_stored(DbSlice& db, const uint8_t* dev_records, size_t n, size_t stride,
                           size_t off_lo, size_t off_hi, uint32_t max_serial, hipStream_t s) const;
  // n u16 rows already on the device -> HBM limbs at `index` (stream-ordered)
  void load_records_device(DbSlice& db, size_t index, size_t n, const uint16_t* dev_rows,
                           hipStream_t s) const;
  // copies rows [src, src + n) to [dst, dst + n) (multiples of 16, no overlap)
  void replicate_records(DbSlice& db, size_t src, size_t n, size_t dst, hipStream_t s) const;
  // prefetch_subset (share_db.rs:344-407): rows ids[0..n) (device u32) of
  // `src` to rows 0..n of `dst` (a subset slice, e.g. alloc_db(DB_CHUNK_SIZE))
  void prefetch_subset(const DbSlice& src, const uint32_t* ids, size_t n, DbSlice& dst,
                       hipStream_t s) const;
  // preprocess_query + htod + query_sums: rows x code_length u16 (already rotated
  // and Lagrange-preprocessed, like the reference's CompactQuery).
  std::unique_ptr<QuerySet> upload_query(const uint16_t* host_rows, size_t n_rows, hipStream_t s) const;
  std::unique_ptr<QuerySet> upload_query_device(const uint16_t* dev_rows, size_t n_rows, hipStream_t s) const;
  // Raw (not preprocessed, not rotated) query shares of this party: batch x
  // code_length u16 on the host. Lagrange factor, basis B and the 31 rotations
  // are applied on the GPU; rows are padded to query_length.
  std::unique_ptr<QuerySet> upload_query_shares(const uint16_t* host_shares, size_t batch,
                                                hipStream_t s) const;
  // Same with exactly `rows` query rows (batch * 31 <= rows; 0 = query_length),
  // e.g. several batches' queries packed back to back for one launch.
  std::unique_ptr<QuerySet> upload_query_shares_rows(const uint16_t* host_shares, size_t batch,
                                                     size_t rows, hipStream_t s) const;
  // Queries of the Mirror orientation (server/mod.rs:156-176): query rows from
  // the mirrored host_shares (mirrored_code for 12,800-coefficient engines,
  // mirrored_mask otherwise; galois_engine.rs:218-224,518-545), the DB form
  // from insert_shares (batch x code_length, this eye's raw shares).
  std::unique_ptr<QuerySet> upload_query_shares_mirrored(const uint16_t* host_shares,
                                                         const uint16_t* insert_shares,
                                                         size_t batch, size_t rows,
                                                         hipStream_t s) const;
  // Refills a query set made by upload_query_shares[_rows] with new raw
  // shares (batch queries, n_rows rows, 0 = query_length), reusing its
  // buffers: no device allocation or free, so no implicit device-wide
  // synchronisation. n_rows and batch must fit what the set was made with.
  void refill_query_shares(QuerySet& q, const uint16_t* host_shares, size_t batch, size_t n_rows,
                           hipStream_t s) const;
  // The same without synchronising: stream-ordered after the work already
  // queued on `s` (which may still read the set) and before work queued
  // later. Every asynchronous refill of this engine must go on one stream
  // (the staging buffers are reused in stream order), and host_shares
  // (pinned host memory) must stay unchanged until the stream has run the
  // copy.
  // with_db_form = false skips the set's DB form (the batch's irises as DB
  // rows, read only by write_at_index and the batch self-comparison), which
  // then goes stale.
  void refill_query_shares_async(QuerySet& q, const uint16_t* host_shares, size_t batch,
                                 size_t n_rows, hipStream_t s, bool with_db_form = true) const;
  // dot + dot_reduce_and_multiply (fused): DB rows [offset, offset + chunk)
  // against all query rows; chunk is the padded chunk size (multiple of 64).
  // results land at results() + out_offset (elements): several chunks of a
  // small query set can sit side by side in the result buffer
  void dot_reduce(const QuerySet& q, const DbSlice& db, size_t chunk, size_t offset,
                  uint16_t multiplier, hipStream_t s, size_t out_offset = 0);
  // dot_reduce for n engines of one device at once (engine i: query q[i], DB
  // db[i], multiplier mult[i], the same chunk and offset), results and masks
  // those of the n calls in order. When every call takes the rotation-aware
  // 4-query kernel they run as one launch (one kernel boundary instead of n);
  // otherwise as n launches. Timed on eng[0]'s timer when it is enabled.
  static void dot_reduce_joint(ShareDB* const eng[], const QuerySet* const q[], const DbSlice* const db[],
                               const uint16_t mult[], int n, size_t chunk, size_t offset, hipStream_t s);
  // Consecutive dot_reduce calls in one launch: DB rows [offset, offset +
  // total) as chunks of `chunk` rows (the last one total mod chunk wide when
  // that is not zero; a multiple of 64), results side by side from results() +
  // out_offset, each chunk with the layout and masks its own dot_reduce call
  // would give (chunk c's mask right after chunk c - 1's: no reshare between).
  void dot_reduce_chunks(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                         size_t chunk, uint16_t multiplier, hipStream_t s, size_t out_offset = 0);
  // Multi-chunk single launch: rows [offset, offset + total) in chunks of
  // `chunk` rows, results written chunk after chunk into `out`.
  void dot_reduce_range(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                        size_t chunk, uint16_t multiplier, uint16_t* out, hipStream_t s);
  // Several batches packed in one query set (rows[i] rows each, back to back)
  // in one launch; batch i's results go to outs[i] with the layout and mask
  // counters its own dot_reduce_range would have used, in order.
  // ctr_first (optional) gives each batch's first ChaCha block explicitly and
  // leaves the engine counter alone (for plans that interleave batches).
  void dot_reduce_packed(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                         size_t chunk, uint16_t multiplier, const size_t* rows,
                         uint16_t* const* outs, int n_batches, const uint64_t* ctr_first,
                         hipStream_t s);
  // The launch arguments dot_reduce_packed would use with GEMM variant
  // `variant` (no launch; the counter stays): for persistent launches that
  // cover several packs. *ctr_end (optional) = the counter after the range.
  GemmArgs packed_args(const QuerySet& q, const DbSlice& db, size_t offset, size_t total, size_t chunk,
                       uint16_t multiplier, const size_t* rows, uint16_t* const* outs, int n_batches,
                       const uint64_t* ctr_first, int variant, uint64_t* ctr_end = nullptr) const;
  int device() const { return device_; }
  void reshare_results(size_t chunk, hipStream_t s);  // collective
  static void reshare_trio(ShareDB* const d[3], size_t chunk, hipStream_t const s[3]);
  uint16_t* results() const { return slot_ ? results_alt_.ptr : results_.ptr; }
  // Result buffer of the following dot_reduce calls: 0 or 1 (the second is
  // allocated on first use), so a chunk's GEMM can run while the previous
  // chunk's circuits still read the other one.
  void set_slot(int slot);
  // The previous party's result share. The reshare hands it over as a view of
  // that party's own result buffer (no copy; valid until that party's next
  // dot_reduce, which the three-party circuit launches order after their
  // reads), or, with peer_copy set on any of the three engines, as a copy.
  uint16_t* results_peer() const {
    return (peer_copy || !peer_view_) ? results_peer_.ptr : peer_view_;
  }
  bool peer_copy = false;
  size_t results_capacity() const { return results_.n; }
  int peer_id() const { return peer_id_; }
  int code_length() const { return code_length_; }
  size_t query_length() const { return query_length_; }
  uint64_t counter() const { return counter_; }
  void set_counter(uint64_t c) { counter_ = c; }
  bool is_remote() const { return ex_ != nullptr; }
  // ChaCha blocks a remote engine's dot_reduce_range consumes (mask + OTP per chunk)
  static uint64_t range_counter_advance(size_t total, size_t chunk, size_t rows);
  // GEMM tile variant; kAutoVariant picks the rotation-aware kernel for query
  // sets with rotation structure and the best generic one otherwise
  static constexpr int kAutoVariant = -1;
  int variant = kAutoVariant;
  int resolve_variant(const QuerySet& q) const;
  // share GEMM launches (every dot_reduce form), when enabled
  LaunchTimer timer;

 private:
  void launch(const GemmArgs& a, int variant, hipStream_t s);
  // dot_reduce's checks, arguments and counter bookkeeping, without the launch
  GemmArgs dot_reduce_args(const QuerySet& q, const DbSlice& db, size_t chunk, size_t offset,
                           uint16_t multiplier, size_t out_offset, int* variant);
  std::unique_ptr<QuerySet> upload_shares_impl(const uint16_t* host_shares,
                                               const uint16_t* insert_shares, size_t batch,
                                               size_t rows, int mirror, hipStream_t s) const;
  int device_, peer_id_, code_length_, tag_;
  size_t query_length_;
  ChaChaKey seed_own_, seed_prev_;
  uint64_t nonce_ = 0, counter_ = 0;
  size_t last_rows_;  // query rows of the last dot_reduce (the reshare length)
  Exchange* ex_;
  DevBuf<uint16_t> results_, results_peer_, results_alt_;
  // query staging of refill_query_shares (raw shares, prepared rows)
  mutable DevBuf<uint16_t> qraw_, qrows_;
  uint16_t* peer_view_ = nullptr;
  int slot_ = 0;
  Event ev_send_done_;
  friend struct ShareDBTrio;
};

// Stored-record layout of the DB chunks (iris-mpc-store/src/s3_importer.rs:
// 13-87, written by iris-mpc-db-exporter/src/converter/binary.go:59-152):
// BE u32 serial id; left code, left mask, right code, right mask, each as its
// low limb bytes ("odd") then its high limb bytes ("even"), biased by -128;
// BE u16 version id.
constexpr size_t kStoredRecordBytes = 4 + 2 * (2 * kCodeLen) + 2 * (2 * kMaskLen) + 2;
struct StoredField {
  size_t lo, hi;
};
constexpr StoredField kStoredLeftCode{4, 4 + kCodeLen};
constexpr StoredField kStoredLeftMask{4 + 2 * kCodeLen, 4 + 2 * kCodeLen + kMaskLen};
constexpr StoredField kStoredRightCode{4 + 2 * kCodeLen + 2 * kMaskLen, 4 + 3 * kCodeLen + 2 * kMaskLen};
constexpr StoredField kStoredRightMask{4 + 4 * kCodeLen + 2 * kMaskLen,
                                       4 + 4 * kCodeLen + 3 * kMaskLen};

// InMemoryStore::load_single_record_from_s3 (inmemory_store.rs:55-105) over n
// stored records in host memory, loader.rs:196-233 placement: serial id s ->
// row s - 1 of all four slices; ids above max_serial (0 = no cap) skipped; id
// 0, an id repeated within the call or a row beyond a slice is an error
// (checked before anything is written). Returns the records loaded; *row_end
// = largest loaded row + 1 (0 if none).
size_t load_stored_records(const ShareDB& codes, const ShareDB& masks, DbSlice& left_code,
                           DbSlice& left_mask, DbSlice& right_code, DbSlice& right_mask,
                           const uint8_t* records, size_t n, uint32_t max_serial,
                           size_t* row_end, hipStream_t s);

class Circuits {
 public:
  Circuits(int device, int peer_id, size_t input_size, size_t alloc_size,
           const ChaChaKey& seed_own, const ChaChaKey& seed_prev, Exchange* ex, int tag);
  void set_chunk_size(size_t words);
  size_t chunk_size() const { return chunk_words_; }
  // compare_threshold_masked_many: collective over the three parties.
  // fork = false: the party streams are not made to wait for the circuits
  // (work that reads the results goes on exchange()->protocol_stream())
  void compare_threshold_masked_many(const uint16_t* code_a, const uint16_t* code_b,
                                     const uint16_t* mask_a, const uint16_t* mask_b, size_t len,
                                     uint32_t threshold_a, hipStream_t s, bool fork = true);
  // Runs the three parties' compare in one call (single host thread).
  static void compare_trio(Circuits* const c[3], const uint16_t* const code_a[3],
                           const uint16_t* const code_b[3], const uint16_t* const mask_a[3],
                           const uint16_t* const mask_b[3], size_t len, uint32_t threshold_a,
                           hipStream_t const s[3], bool fork = true);
  Exchange* exchange() const { return ex_; }
  // open exchange (actor.rs:2928-2997): own result b -> next, prev's b -> c.
  void open_exchange(hipStream_t s);  // collective
  static void open_exchange_trio(Circuits* const c[3], hipStream_t const s[3]);
  const uint64_t* result_a() const { return res_a_.ptr; }
  const uint64_t* result_b() const { return res_b_.ptr; }
  // the previous party's result b, handed over as a view of its buffer (valid
  // until the next compare, whose launches wait for this party's stream)
  const uint64_t* result_c() const { return open_view_ ? open_view_ : open_c_.ptr; }
  uint64_t counter() const { return counter_; }
  void set_counter(uint64_t c) { counter_ = c; }
  int peer_id() const { return peer_id_; }
  uint32_t* lifted_dbg_a = nullptr;  // optional debug capture of the lifted mask
  uint32_t* lifted_dbg_b = nullptr;
  // the circuit launches of a compare (lift, injection, MSB adder: one timed
  // interval on the protocol stream, kept by party 0's engine), when enabled
  LaunchTimer timer;

 private:
  CircuitBufs bufs(const uint16_t* ca, const uint16_t* cb, const uint16_t* ma,
                   const uint16_t* mb) const;
  int device_, peer_id_, tag_;
  size_t alloc_words_, chunk_words_;
  ChaChaKey seed_own_, seed_prev_;
  uint64_t nonce_ = 0, counter_ = 0;
  Exchange* ex_;
  DevBuf<uint64_t> c_a_, c_b_, res_a_, res_b_, open_c_;
  DevBuf<uint16_t> corr_a_;
  const uint64_t* open_view_ = nullptr;
};

// DistanceComparator analogue (dot/distance_comparator.rs).
class DistanceComparator {
 public:
  DistanceComparator(int device, size_t query_length, size_t max_db_size);
  // Bitmap of size max_db * batch bits (prepare_db_match_list).
  std::unique_ptr<DevBuf<uint64_t>> prepare_db_match_list(size_t db_size) const;
  // rows: query rows of the circuit result (0 = query_length); slot_of (device,
  // optional): batch query j's bits go to bitmap row slot_of[j] instead of j
  // dots (optional): the chunk's dot-product shares, for the match-distance
  // capture when one is set
  void open_results(const uint64_t* a, const uint64_t* b, const uint64_t* c, uint64_t* bitmap,
                    size_t chunk_len, size_t real_chunk_len, size_t offset, size_t total_db_len,
                    size_t batch_size, hipStream_t s, size_t rows = 0,
                    const uint32_t* slot_of = nullptr, const DotViews* dots = nullptr) const;
  // Match-distance capture of the following opens (openResults' anonymized-
  // statistics buffers, dot/kernel.cu:73-106; server/anon_stats.rs): device
  // buffers of `capacity` entries and a counter; targets (host, db * max_batch
  // + query) are also captured when not matching. capacity 0 switches it off.
  void set_anon_capture(const AnonCapture& c, const uint64_t* host_targets, size_t n_targets);
  bool anon_capture_enabled() const { return anon_.enabled(); }
  // join_db_matches (AND policy) and join_db_matches_with_bitmaps (OR policy).
  void join_db_matches(const uint64_t* left, const uint64_t* right, uint32_t* final_results,
                       size_t db_size, const uint64_t* or_policy, size_t max_db_size,
                       hipStream_t s);
  // join_batch_matches (distance_comparator.rs:731-746, mergeBatchResults
  // dot/kernel.cu:352-386) over batch self-comparison bitmaps: bit
  // row_bits * q + j = query q matches query j's unrotated iris. For j < q
  // matching in both eyes: final_results[q] = min(., 0xFFFFFFFE) and
  // 0xFFFFFFFF - j joins q's match list.
  void join_batch_matches(const uint64_t* left, const uint64_t* right, uint32_t* final_results,
                          size_t batch_size, size_t row_bits, hipStream_t s);
  // get_partial_results (distance_comparator.rs:545-648) plus the OR-rule and
  // reauth indices the actor appends (actor.rs:946-970): the ascending list of
  // DB ids d < db_size set in any batch row of `bitmap` (rows row_bits bits
  // apart; 0 = db_size) or listed in extra_host, kept on the device
  // (partial_ids()). Returns its length (synchronises the stream).
  size_t get_partial_results(const uint64_t* bitmap, size_t db_size, size_t batch_size,
                             const uint32_t* extra_host, size_t n_extra, hipStream_t s,
                             size_t row_bits = 0);
  const uint32_t* partial_ids() const { return ids_.ptr; }
  // open_results_with_index_mapping (distance_comparator.rs:281-342): subset
  // chunk column j is DB row index_map[j] (device); matches also go to the
  // rotation capture (partial_results_*, read by fetch_partial_rotations)
  // With dots set and bitmap == nullptr: the anonymized-statistics pass of
  // the subset (storeAnonStatsWithIndexMapping, dot/kernel.cu:153-210), which
  // only captures distances.
  void open_results_mapped(const uint64_t* a, const uint64_t* b, const uint64_t* c,
                           uint64_t* bitmap, size_t chunk_len, size_t real_chunk_len,
                           size_t total_db_len, size_t batch_size, const uint32_t* index_map,
                           hipStream_t s, size_t rows = 0, const uint32_t* slot_of = nullptr,
                           const DotViews* dots = nullptr);
  // get_partial_results_with_rotations (distance_comparator.rs:650-711):
  // copies up to `capacity` captured (query, db id, rotation) triples to the
  // host, resets the capture and returns the number of matches seen (may
  // exceed the capture's size, 256 * query_length)
  size_t fetch_partial_rotations(uint32_t* query, uint32_t* db, int8_t* rot, size_t capacity,
                                 hipStream_t s);
  size_t rotation_capture_size() const { return cap_q_.n; }
  uint32_t* match_counters() const { return match_counters_.ptr; }
  uint32_t* all_matches() const { return all_matches_.ptr; }
  uint32_t* match_counters_left() const { return mcl_.ptr; }
  uint32_t* match_counters_right() const { return mcr_.ptr; }
  uint32_t* partial_results_left() const { return prl_.ptr; }
  uint32_t* partial_results_right() const { return prr_.ptr; }
  size_t query_length() const { return query_length_; }
  size_t max_db_size() const { return max_db_size_; }
  void reset(hipStream_t s);

 private:
  int device_;
  size_t query_length_, max_db_size_;
  DevBuf<uint32_t> match_counters_, all_matches_, mcl_, mcr_, prl_, prr_;
  DevBuf<uint64_t> cand_;
  DevBuf<uint32_t> tiles_, ids_, count_, extra_;
  DevBuf<uint32_t> cap_counter_, cap_q_, cap_db_;
  DevBuf<int8_t> cap_rot_;
  AnonCapture anon_;
  DevBuf<uint64_t> anon_targets_;
};

}  // namespace irismpc
