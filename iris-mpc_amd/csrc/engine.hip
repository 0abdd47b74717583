// Host engines: Exchange, ShareDB, Circuits, DistanceComparator.
#include "engine.hpp"
#include "comparator.hpp"
#include "query_prep.hpp"

#include <algorithm>
#include <vector>
#include <chrono>
#include <cstring>

namespace irismpc {

namespace {
struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int d) {
    IRIS_HIP_CHECK(hipGetDevice(&prev));
    if (prev != d) IRIS_HIP_CHECK(hipSetDevice(d));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

uint64_t op_id(const char* name, int tag) {
  uint64_t h = 1469598103934665603ull;
  for (const char* p = name; *p; ++p) h = (h ^ (uint8_t)*p) * 1099511628211ull;
  return h ^ ((uint64_t)tag << 48);
}
}  // namespace

// ====================================================================== LaunchTimer
LaunchTimer::~LaunchTimer() {
  for (auto& p : ev_) {
    if (p.a) (void)hipEventDestroy(p.a);
    if (p.b) (void)hipEventDestroy(p.b);
  }
}

void LaunchTimer::begin(hipStream_t s) {
  if (used_ == ev_.size()) {
    Pair p;
    IRIS_HIP_CHECK(hipEventCreate(&p.a));
    IRIS_HIP_CHECK(hipEventCreate(&p.b));
    ev_.push_back(p);
  }
  IRIS_HIP_CHECK(hipEventRecord(ev_[used_].a, s));
}

void LaunchTimer::end(hipStream_t s) {
  IRIS_HIP_CHECK(hipEventRecord(ev_[used_].b, s));
  used_++;
}

size_t LaunchTimer::collect(double* total_ms) {
  double t = 0;
  for (size_t i = 0; i < used_; i++) {
    IRIS_HIP_CHECK(hipEventSynchronize(ev_[i].b));
    float ms = 0;
    IRIS_HIP_CHECK(hipEventElapsedTime(&ms, ev_[i].a, ev_[i].b));
    t += ms;
  }
  const size_t n = used_;
  used_ = 0;
  if (total_ms) *total_ms = t;
  return n;
}

// ====================================================================== Exchange
Exchange::Exchange(int device, double timeout_s) : device_(device), timeout_s_(timeout_s) {
  DeviceGuard g(device_);
  IRIS_HIP_CHECK(hipStreamCreateWithFlags(&protocol_, hipStreamNonBlocking));
}

Exchange::~Exchange() {
  if (protocol_) (void)hipStreamDestroy(protocol_);
}

void Exchange::collective(int party, uint64_t op, void* arg,
                          const std::function<void(void* const*)>& fn) {
  IRIS_REQUIRE(party >= 0 && party < 3, "party id must be 0, 1 or 2");
  std::unique_lock<std::mutex> lk(mu_);
  const uint64_t gen = generation_;
  IRIS_REQUIRE(args_[party] == nullptr || arrived_ == 0, "party joined the same collective twice");
  args_[party] = arg;
  ops_[party] = op;
  if (++arrived_ == 3) {
    std::exception_ptr err;
    if (!(ops_[0] == ops_[1] && ops_[1] == ops_[2])) {
      err = std::make_exception_ptr(Error("collective mismatch: parties called different operations"));
    } else {
      try {
        fn(args_);
      } catch (...) {
        err = std::current_exception();
      }
    }
    for (auto& a : args_) a = nullptr;
    arrived_ = 0;
    error_ = err;
    error_gen_ = gen;
    ++generation_;
    cv_.notify_all();
    if (err) std::rethrow_exception(err);
    return;
  }
  const bool done = cv_.wait_for(lk, std::chrono::duration<double>(timeout_s_),
                                 [&] { return generation_ != gen; });
  if (!done) {
    --arrived_;
    args_[party] = nullptr;
    throw Error("exchange timeout: a peer party did not join the collective");
  }
  if (error_gen_ == gen && error_) std::rethrow_exception(error_);
}

// True for the parties whose stream is not the protocol stream and not an
// earlier party's: one event wait per distinct stream (a wait on a stream's
// own work, or a second wait on the same work, would only cost a marker).
static bool own_party_stream(hipStream_t const s[3], int p, hipStream_t protocol) {
  if (s[p] == protocol) return false;
  for (int q = 0; q < p; q++)
    if (s[q] == s[p]) return false;
  return true;
}

void Exchange::join_streams(hipStream_t const s[3]) {
  for (int p = 0; p < 3; p++) {
    if (!own_party_stream(s, p, protocol_)) continue;
    IRIS_HIP_CHECK(hipEventRecord(ev_party_[p].e, s[p]));
    IRIS_HIP_CHECK(hipStreamWaitEvent(protocol_, ev_party_[p].e, 0));
  }
}

void Exchange::fork_streams(hipStream_t const s[3]) {
  IRIS_HIP_CHECK(hipEventRecord(ev_proto_.e, protocol_));
  for (int p = 0; p < 3; p++)
    if (own_party_stream(s, p, protocol_)) IRIS_HIP_CHECK(hipStreamWaitEvent(s[p], ev_proto_.e, 0));
}

// ====================================================================== ShareDB
ShareDB::ShareDB(int device, int peer_id, size_t max_db_length, size_t query_length,
                 int code_length, const ChaChaKey& seed_own, const ChaChaKey& seed_prev,
                 Exchange* ex, int tag)
    : device_(device),
      peer_id_(peer_id),
      code_length_(code_length),
      tag_(tag),
      query_length_(query_length),
      seed_own_(seed_own),
      seed_prev_(seed_prev),
      last_rows_(query_length),
      ex_(ex) {
  IRIS_REQUIRE(peer_id >= 0 && peer_id < 3, "peer_id must be 0..2");
  IRIS_REQUIRE(code_length % kKTile == 0 && code_length % 800 == 0,
               "code_length must be a multiple of 800 (rows of 200 columns x 4 coefs)");
  IRIS_REQUIRE(query_length > 0 && max_db_length > 0, "empty engine");
  DeviceGuard g(device_);
  const size_t rl = round_up(max_db_length * query_length, 64);
  results_.alloc(rl);
  results_peer_.alloc(rl);
}

ShareDB::~ShareDB() = default;

void ShareDB::set_slot(int slot) {
  IRIS_REQUIRE(slot == 0 || slot == 1, "result slot must be 0 or 1");
  if (slot == 1 && !results_alt_.ptr) {
    DeviceGuard g(device_);
    results_alt_.alloc(results_.n);
  }
  slot_ = slot;
}

std::unique_ptr<DbSlice> ShareDB::alloc_db(size_t max_db_length) const {
  DeviceGuard g(device_);
  auto db = std::make_unique<DbSlice>();
  db->code_length = code_length_;
  // capacity padded so every GEMM tile (up to 256 rows from any 16-aligned
  // start) stays inside the allocation
  db->capacity = round_up(max_db_length, kMaxBD) + kMaxBD;
  db->limbs.alloc(db->capacity * (size_t)code_length_ * 2);
  db->sum0.alloc(db->capacity);
  db->sum1.alloc(db->capacity);
  return db;
}

void ShareDB::load_records_u16(DbSlice& db, size_t index, size_t n, const uint16_t* host,
                               hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(index + n <= db.capacity - kMaxBD, "record index beyond DB capacity");
  const size_t K = (size_t)code_length_;
  const size_t batch = std::max<size_t>(1, (64ull << 20) / (K * 2));
  if (db.staging.n < batch * K) db.staging.alloc(batch * K, false);
  for (size_t done = 0; done < n; done += batch) {
    const size_t m = std::min(batch, n - done);
    IRIS_HIP_CHECK(hipMemcpyAsync(db.staging.ptr, host + done * K, m * K * 2, hipMemcpyHostToDevice, s));
    launch_tile_rows(db.staging.ptr, m, code_length_, db.limbs.ptr, index + done, db.sum0.ptr,
                     db.sum1.ptr, s);
    IRIS_HIP_CHECK(hipStreamSynchronize(s));  // staging reuse
  }
}

void ShareDB::load_records_stored(DbSlice& db, const uint8_t* dev_records, size_t n,
                                  size_t stride, size_t off_lo, size_t off_hi,
                                  uint32_t max_serial, hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(db.code_length == code_length_, "DB slice belongs to another code length");
  IRIS_REQUIRE(off_lo + (size_t)code_length_ <= stride && off_hi + (size_t)code_length_ <= stride,
               "field beyond the record");
  launch_tile_records(dev_records, n, stride, off_lo, off_hi, code_length_, max_serial,
                      db.limbs.ptr, db.sum0.ptr, db.sum1.ptr, s);
}

size_t load_stored_records(const ShareDB& codes, const ShareDB& masks, DbSlice& left_code,
                           DbSlice& left_mask, DbSlice& right_code, DbSlice& right_mask,
                           const uint8_t* records, size_t n, uint32_t max_serial,
                           size_t* row_end, hipStream_t s) {
  IRIS_REQUIRE(codes.code_length() == kCodeLen && masks.code_length() == kMaskLen,
               "stored records hold 12,800-coefficient codes and 6,400-coefficient masks");
  if (max_serial == 0) max_serial = 0xFFFFFFFFu;
  size_t cap = std::min({left_code.capacity, left_mask.capacity, right_code.capacity,
                         right_mask.capacity}) - kMaxBD;
  std::vector<uint32_t> ids;
  size_t loaded = 0, end = 0;
  for (size_t r = 0; r < n; r++) {
    const uint8_t* b = records + r * kStoredRecordBytes;
    const uint32_t id = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
    IRIS_REQUIRE(id != 0, "stored record with serial id 0");
    if (id > max_serial) continue;
    IRIS_REQUIRE((size_t)id - 1 < cap, "serial id beyond the DB slices");
    ids.push_back(id);
    loaded++;
    end = std::max(end, (size_t)id);
  }
  std::sort(ids.begin(), ids.end());
  IRIS_REQUIRE(std::adjacent_find(ids.begin(), ids.end()) == ids.end(),
               "serial id repeated within the records");
  const size_t window = std::min<size_t>(n, 512);
  DevBuf<uint8_t> stage(window * kStoredRecordBytes, false);
  for (size_t r0 = 0; r0 < n; r0 += window) {
    const size_t m = std::min(window, n - r0);
    IRIS_HIP_CHECK(hipMemcpyAsync(stage.ptr, records + r0 * kStoredRecordBytes,
                                  m * kStoredRecordBytes, hipMemcpyHostToDevice, s));
    codes.load_records_stored(left_code, stage.ptr, m, kStoredRecordBytes, kStoredLeftCode.lo,
                              kStoredLeftCode.hi, max_serial, s);
    masks.load_records_stored(left_mask, stage.ptr, m, kStoredRecordBytes, kStoredLeftMask.lo,
                              kStoredLeftMask.hi, max_serial, s);
    codes.load_records_stored(right_code, stage.ptr, m, kStoredRecordBytes, kStoredRightCode.lo,
                              kStoredRightCode.hi, max_serial, s);
    masks.load_records_stored(right_mask, stage.ptr, m, kStoredRecordBytes, kStoredRightMask.lo,
                              kStoredRightMask.hi, max_serial, s);
    IRIS_HIP_CHECK(hipStreamSynchronize(s));  // staging reuse
  }
  if (row_end) *row_end = end;
  return loaded;
}

void ShareDB::load_records_device(DbSlice& db, size_t index, size_t n, const uint16_t* dev_rows,
                                  hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(index + n <= db.capacity - kMaxBD, "record index beyond DB capacity");
  launch_tile_rows(dev_rows, n, code_length_, db.limbs.ptr, index, db.sum0.ptr, db.sum1.ptr, s);
}

void ShareDB::replicate_records(DbSlice& db, size_t src, size_t n, size_t dst, hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(src % 16 == 0 && dst % 16 == 0 && n % 16 == 0, "replicate needs 16-row groups");
  IRIS_REQUIRE(src + n <= dst || dst + n <= src, "replicate regions overlap");
  IRIS_REQUIRE(std::max(src, dst) + n <= db.capacity - kMaxBD, "replicate beyond DB capacity");
  // a 16-row group is (K / 64) * 2 fragments of 1 KiB, contiguous per group
  const size_t group_bytes = (size_t)code_length_ / kKTile * 2 * kFragBytes;
  IRIS_HIP_CHECK(hipMemcpyAsync(db.limbs.ptr + dst / 16 * group_bytes, db.limbs.ptr + src / 16 * group_bytes,
                                n / 16 * group_bytes, hipMemcpyDeviceToDevice, s));
  IRIS_HIP_CHECK(hipMemcpyAsync(db.sum0.ptr + dst, db.sum0.ptr + src, n * 4, hipMemcpyDeviceToDevice, s));
  IRIS_HIP_CHECK(hipMemcpyAsync(db.sum1.ptr + dst, db.sum1.ptr + src, n * 4, hipMemcpyDeviceToDevice, s));
}

void ShareDB::prefetch_subset(const DbSlice& src, const uint32_t* ids, size_t n, DbSlice& dst,
                              hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(src.code_length == code_length_ && dst.code_length == code_length_,
               "subset and DB slices belong to another code length");
  IRIS_REQUIRE(n <= dst.capacity - kMaxBD, "subset larger than the subset slice");
  launch_gather_rows(src.limbs.ptr, src.sum0.ptr, src.sum1.ptr, src.capacity - kMaxBD, ids, n,
                     dst.limbs.ptr, dst.sum0.ptr, dst.sum1.ptr, code_length_, s);
}

void ShareDB::load_records_limbs(DbSlice& db, size_t index, size_t n, const uint8_t* l0,
                                 const uint8_t* l1, hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(index + n <= db.capacity - kMaxBD, "record index beyond DB capacity");
  const size_t K = (size_t)code_length_;
  const size_t batch = std::max<size_t>(1, (64ull << 20) / (K * 2));
  if (db.staging.n < batch * K) db.staging.alloc(batch * K, false);
  uint8_t* s0 = reinterpret_cast<uint8_t*>(db.staging.ptr);
  uint8_t* s1 = s0 + batch * K;
  for (size_t done = 0; done < n; done += batch) {
    const size_t m = std::min(batch, n - done);
    IRIS_HIP_CHECK(hipMemcpyAsync(s0, l0 + done * K, m * K, hipMemcpyHostToDevice, s));
    IRIS_HIP_CHECK(hipMemcpyAsync(s1, l1 + done * K, m * K, hipMemcpyHostToDevice, s));
    launch_tile_limbs(s0, s1, m, code_length_, db.limbs.ptr, index + done, db.sum0.ptr, db.sum1.ptr, s);
    IRIS_HIP_CHECK(hipStreamSynchronize(s));
  }
}

void ShareDB::write_at_index(DbSlice& db, const QuerySet& q, size_t src_index, size_t dst_index,
                             hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(q.code_length == code_length_, "query/engine code length mismatch");
  IRIS_REQUIRE(dst_index < db.capacity - kMaxBD, "dst_index beyond DB capacity");
  if (q.insert) {  // the query's DB form
    IRIS_REQUIRE(src_index < (size_t)q.batch, "src_index beyond query batch");
    const DbSlice& in = *q.insert;
    launch_copy_row(in.limbs.ptr, in.sum0.ptr, in.sum1.ptr, src_index, db.limbs.ptr, db.sum0.ptr,
                    db.sum1.ptr, dst_index, code_length_, s);
    return;
  }
  // rows given by the caller: the centre rotation (rotation 15) of query
  // src_index, as the reference copies it from its insert-form rows
  const size_t src_row = src_index * kRotations + 15;
  IRIS_REQUIRE(src_row < (size_t)q.rows, "src_index beyond query batch");
  launch_copy_row(q.limbs.ptr, q.sum0.ptr, q.sum1.ptr, src_row, db.limbs.ptr, db.sum0.ptr,
                  db.sum1.ptr, dst_index, code_length_, s);
}

std::unique_ptr<QuerySet> ShareDB::upload_query_device(const uint16_t* dev_rows, size_t n_rows,
                                                       hipStream_t s) const {
  DeviceGuard g(device_);
  IRIS_REQUIRE(n_rows == query_length_, "query rows must equal the engine's query_length");
  auto q = std::make_unique<QuerySet>();
  q->code_length = code_length_;
  q->rows = (int)n_rows;
  const size_t padded = round_up(n_rows, kMaxBQ);
  q->limbs.alloc(padded * (size_t)code_length_ * 2);
  q->sum0.alloc(padded);
  q->sum1.alloc(padded);
  launch_tile_rows(dev_rows, n_rows, code_length_, q->limbs.ptr, 0, q->sum0.ptr, q->sum1.ptr, s);
  return q;
}

std::unique_ptr<QuerySet> ShareDB::upload_query(const uint16_t* host_rows, size_t n_rows,
                                                hipStream_t s) const {
  DeviceGuard g(device_);
  DevBuf<uint16_t> tmp(n_rows * (size_t)code_length_, false);
  IRIS_HIP_CHECK(hipMemcpyAsync(tmp.ptr, host_rows, tmp.n * 2, hipMemcpyHostToDevice, s));
  auto q = upload_query_device(tmp.ptr, n_rows, s);
  IRIS_HIP_CHECK(hipStreamSynchronize(s));
  return q;
}

std::unique_ptr<QuerySet> ShareDB::upload_query_shares(const uint16_t* host_shares, size_t batch,
                                                       hipStream_t s) const {
  return upload_query_shares_rows(host_shares, batch, query_length_, s);
}

std::unique_ptr<QuerySet> ShareDB::upload_query_shares_rows(const uint16_t* host_shares, size_t batch,
                                                            size_t n_rows, hipStream_t s) const {
  return upload_shares_impl(host_shares, nullptr, batch, n_rows, 0, s);
}

std::unique_ptr<QuerySet> ShareDB::upload_query_shares_mirrored(const uint16_t* host_shares,
                                                                const uint16_t* insert_shares,
                                                                size_t batch, size_t n_rows,
                                                                hipStream_t s) const {
  IRIS_REQUIRE(insert_shares, "mirrored queries need this eye's raw shares for the DB form");
  return upload_shares_impl(host_shares, insert_shares, batch, n_rows,
                            code_length_ == kCodeLen ? 2 : 1, s);
}

std::unique_ptr<QuerySet> ShareDB::upload_shares_impl(const uint16_t* host_shares,
                                                      const uint16_t* insert_shares, size_t batch,
                                                      size_t n_rows, int mirror,
                                                      hipStream_t s) const {
  DeviceGuard g(device_);
  if (n_rows == 0) n_rows = query_length_;
  IRIS_REQUIRE(batch >= 1 && batch * kRotations <= n_rows, "batch exceeds the query rows / 31");
  const size_t K = (size_t)code_length_;
  DevBuf<uint16_t> raw(batch * K, false), rows(n_rows * K, false);
  IRIS_HIP_CHECK(hipMemcpyAsync(raw.ptr, host_shares, raw.n * 2, hipMemcpyHostToDevice, s));
  launch_query_prep(raw.ptr, (int)batch, code_length_, peer_id_, rows.ptr, (int)n_rows, s, mirror);
  auto q = std::make_unique<QuerySet>();
  q->code_length = code_length_;
  q->rows = (int)n_rows;
  const size_t padded = round_up(n_rows, kMaxBQ);
  q->limbs.alloc(padded * K * 2);
  q->sum0.alloc(padded);
  q->sum1.alloc(padded);
  launch_tile_rows(rows.ptr, n_rows, code_length_, q->limbs.ptr, 0, q->sum0.ptr, q->sum1.ptr, s);
  if (code_length_ % 800 == 0) {
    q->rot.alloc(rot_planes_bytes((int)n_rows, code_length_));
    launch_rot_planes(rows.ptr, (int)n_rows, code_length_, q->rot.ptr, s);
  }
  if (insert_shares) {  // the DB form comes from other shares than the query rows
    IRIS_HIP_CHECK(hipStreamSynchronize(s));  // raw is reused
    IRIS_HIP_CHECK(hipMemcpyAsync(raw.ptr, insert_shares, raw.n * 2, hipMemcpyHostToDevice, s));
  }
  q->batch = (int)batch;
  q->cap_rows = n_rows;
  q->cap_batch = round_up(batch, 512);
  q->insert = alloc_db(q->cap_batch);  // room for a padded self-comparison chunk
  launch_tile_rows(raw.ptr, batch, code_length_, q->insert->limbs.ptr, 0, q->insert->sum0.ptr,
                   q->insert->sum1.ptr, s);
  IRIS_HIP_CHECK(hipStreamSynchronize(s));
  return q;
}

void ShareDB::refill_query_shares(QuerySet& q, const uint16_t* host_shares, size_t batch,
                                  size_t n_rows, hipStream_t s) const {
  refill_query_shares_async(q, host_shares, batch, n_rows, s);
  IRIS_HIP_CHECK(hipStreamSynchronize(s));  // the staging buffers are reused by the next refill
}

void ShareDB::refill_query_shares_async(QuerySet& q, const uint16_t* host_shares, size_t batch,
                                        size_t n_rows, hipStream_t s, bool with_db_form) const {
  DeviceGuard g(device_);
  if (n_rows == 0) n_rows = query_length_;
  IRIS_REQUIRE(q.code_length == code_length_, "query set of another code length");
  IRIS_REQUIRE(q.insert != nullptr && q.cap_rows > 0, "refill needs a set uploaded from raw shares");
  IRIS_REQUIRE(batch >= 1 && batch * kRotations <= n_rows, "batch exceeds the query rows / 31");
  IRIS_REQUIRE(n_rows <= q.cap_rows && batch <= q.cap_batch, "refill larger than the query set");
  const size_t K = (size_t)code_length_;
  if (qraw_.n < batch * K) qraw_.alloc(q.cap_rows / kRotations * K, false);
  if (qrows_.n < n_rows * K) qrows_.alloc(q.cap_rows * K, false);
  IRIS_HIP_CHECK(hipMemcpyAsync(qraw_.ptr, host_shares, batch * K * 2, hipMemcpyHostToDevice, s));
  launch_query_prep(qraw_.ptr, (int)batch, code_length_, peer_id_, qrows_.ptr, (int)n_rows, s, 0);
  launch_tile_rows(qrows_.ptr, n_rows, code_length_, q.limbs.ptr, 0, q.sum0.ptr, q.sum1.ptr, s);
  if (q.rot.ptr) launch_rot_planes(qrows_.ptr, (int)n_rows, code_length_, q.rot.ptr, s);
  if (with_db_form)
    launch_tile_rows(qraw_.ptr, batch, code_length_, q.insert->limbs.ptr, 0, q.insert->sum0.ptr,
                     q.insert->sum1.ptr, s);
  q.rows = (int)n_rows;
  q.batch = (int)batch;
}

void ShareDB::launch(const GemmArgs& a, int variant, hipStream_t s) {
  if (timer.on()) timer.begin(s);
  launch_share_gemm(a, variant, s);
  if (timer.on()) timer.end(s);
}

int ShareDB::resolve_variant(const QuerySet& q) const {
  if (variant == kAutoVariant) {
    if (!q.rot.ptr) return 0;
    if (q.rows <= kRotSmallRows) return kRotSmallVariant;
    // Small sets are bound by DB reads, and the 4-query tile reads the DB once
    // per query tile: where its last tile would be at most half full (5, 6, 9,
    // 10 queries) the two-query tile reads it in fewer passes' worth of time
    // (measured 6, 9, 2.6 and 2.4 % faster over 1M rows; at 3, 4, 7, 8 queries
    // the 4-query tile is 15-17 % faster). Larger sets are MFMA-bound, where
    // the 4-query tile's operand reuse wins.
    const int n = q.rows / kRotations;
    if (q.rows % kRotations == 0 && n <= kRotSmallAlsoMax && (n % 4 == 1 || n % 4 == 2)) return kRotSmallVariant;
    return kRotVariantBase;
  }
  IRIS_REQUIRE(!is_rot_variant(variant) || q.rot.ptr,
               "rotation-aware GEMM variant needs a query uploaded from raw shares");
  return variant;
}

static GemmArgs gemm_args(const QuerySet& q, const DbSlice& db, int variant) {
  GemmArgs a{};
  const GemmTile t = gemm_tile(variant);
  a.q = q.limbs.ptr;
  a.qrot = q.rot.ptr;
  a.db = db.limbs.ptr;
  a.qsum0 = q.sum0.ptr;
  a.qsum1 = q.sum1.ptr;
  a.dsum0 = db.sum0.ptr;
  a.dsum1 = db.sum1.ptr;
  a.n_kt = db.code_length / kKTile;
  a.qr = q.rows;
  a.n_qtiles = (int)ceil_div((size_t)q.rows, (size_t)t.bq);
  return a;
}

GemmArgs ShareDB::dot_reduce_args(const QuerySet& q, const DbSlice& db, size_t chunk, size_t offset,
                                  uint16_t multiplier, size_t out_offset, int* variant_out) {
  IRIS_REQUIRE(q.code_length == code_length_ && db.code_length == code_length_, "code length mismatch");
  IRIS_REQUIRE(q.rows >= 1 && (size_t)q.rows <= query_length_, "query rows exceed query_length");
  IRIS_REQUIRE(chunk > 0 && chunk % 64 == 0, "chunk size must be a positive multiple of 64");
  IRIS_REQUIRE(out_offset + chunk * (size_t)q.rows <= results_.n,
               "chunk larger than the engine's result buffer");
  IRIS_REQUIRE(offset % 16 == 0, "offset must be a multiple of 16");
  IRIS_REQUIRE(offset + chunk <= db.capacity - kMaxBD, "chunk beyond DB capacity");
  const int variant = resolve_variant(q);
  GemmArgs a = gemm_args(q, db, variant);
  const GemmTile t = gemm_tile(variant);
  a.out = results() + out_offset;
  a.out_chunk_stride = 0;
  a.n_dtiles = (int)ceil_div(chunk, (size_t)t.bd);
  a.db_begin = offset;
  a.db_end = offset + chunk;
  a.chunk = round_up(chunk, 256);
  a.ctr_first = counter_;
  a.ctr_per_chunk = 0;
  a.nonce = nonce_;
  a.key_own = seed_own_;
  a.key_prev = seed_prev_;
  a.mult = multiplier;
  a.masked = is_remote() ? 1 : 0;
  // the reference fills the mask keystream (chunk*Qr*2 bytes) only when remote
  last_rows_ = (size_t)q.rows;
  if (is_remote()) counter_ += ceil_div(chunk * last_rows_ * 2, 64);
  *variant_out = variant;
  return a;
}

void ShareDB::dot_reduce(const QuerySet& q, const DbSlice& db, size_t chunk, size_t offset,
                         uint16_t multiplier, hipStream_t s, size_t out_offset) {
  DeviceGuard g(device_);
  int variant;
  const GemmArgs a = dot_reduce_args(q, db, chunk, offset, multiplier, out_offset, &variant);
  launch(a, variant, s);
}

void ShareDB::dot_reduce_joint(ShareDB* const eng[], const QuerySet* const q[], const DbSlice* const db[],
                               const uint16_t mult[], int n, size_t chunk, size_t offset, hipStream_t s) {
  IRIS_REQUIRE(n >= 1 && n <= kMaxHeads, "1 .. 6 engines per joint dot_reduce");
  DeviceGuard g(eng[0]->device_);
  GemmArgs args[kMaxHeads];
  int variant[kMaxHeads];
  bool joint = true;
  for (int i = 0; i < n; i++) {
    IRIS_REQUIRE(eng[i]->device_ == eng[0]->device_, "joint dot_reduce engines share one device");
    args[i] = eng[i]->dot_reduce_args(*q[i], *db[i], chunk, offset, mult[i], 0, &variant[i]);
    joint = joint && variant[i] == kRotVariantBase;
  }
  if (!joint) {
    for (int i = 0; i < n; i++) eng[i]->launch(args[i], variant[i], s);
    return;
  }
  GemmHead heads[kMaxHeads];
  for (int i = 0; i < n; i++) heads[i] = args[i];
  LaunchTimer& timer = eng[0]->timer;
  if (timer.on()) timer.begin(s);
  launch_share_gemm_heads(heads, n, s);
  if (timer.on()) timer.end(s);
}

void ShareDB::dot_reduce_chunks(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                                size_t chunk, uint16_t multiplier, hipStream_t s, size_t out_offset) {
  DeviceGuard g(device_);
  IRIS_REQUIRE(q.code_length == code_length_ && db.code_length == code_length_, "code length mismatch");
  IRIS_REQUIRE(q.rows >= 1 && (size_t)q.rows <= query_length_, "query rows exceed query_length");
  IRIS_REQUIRE(total > 0 && total % 64 == 0 && chunk % 256 == 0, "bad chunk run geometry");
  IRIS_REQUIRE(out_offset + total * (size_t)q.rows <= results_.n,
               "chunks larger than the engine's result buffer");
  IRIS_REQUIRE(offset % 16 == 0 && offset + total <= db.capacity - kMaxBD, "chunks beyond DB capacity");
  const int variant = resolve_variant(q);
  GemmArgs a = gemm_args(q, db, variant);
  const GemmTile t = gemm_tile(variant);
  const size_t qr = (size_t)q.rows;
  a.out = results() + out_offset;
  a.out_chunk_stride = chunk * qr;
  a.n_dtiles = (int)ceil_div(total, (size_t)t.bd);
  a.db_begin = offset;
  a.db_end = offset + total;
  a.chunk = chunk;
  a.ctr_first = counter_;
  a.ctr_per_chunk = ceil_div(chunk * qr * 2, 64);  // mask keystream of a full chunk
  a.nonce = nonce_;
  a.key_own = seed_own_;
  a.key_prev = seed_prev_;
  a.mult = multiplier;
  a.masked = is_remote() ? 1 : 0;
  launch(a, variant, s);
  last_rows_ = qr;
  if (is_remote())
    for (size_t c0 = 0; c0 < total; c0 += chunk) counter_ += ceil_div(std::min(chunk, total - c0) * qr * 2, 64);
}

void ShareDB::dot_reduce_range(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                               size_t chunk, uint16_t multiplier, uint16_t* out, hipStream_t s) {
  DeviceGuard g(device_);
  IRIS_REQUIRE(q.code_length == code_length_ && db.code_length == code_length_, "code length mismatch");
  IRIS_REQUIRE(total > 0 && total % 64 == 0 && chunk % 256 == 0, "bad range geometry");
  IRIS_REQUIRE(offset % 16 == 0 && offset + total <= db.capacity - kMaxBD, "range beyond DB");
  const int variant = resolve_variant(q);
  GemmArgs a = gemm_args(q, db, variant);
  const GemmTile t = gemm_tile(variant);
  const size_t qr = (size_t)q.rows;
  a.out = out;
  a.out_chunk_stride = chunk * qr;
  a.n_dtiles = (int)ceil_div(total, (size_t)t.bd);
  a.db_begin = offset;
  a.db_end = offset + total;
  a.chunk = chunk;
  a.ctr_first = counter_;
  // per full chunk: mask keystream + reshare OTP keystream
  a.ctr_per_chunk = 2 * ceil_div(chunk * qr * 2, 64);
  a.nonce = nonce_;
  a.key_own = seed_own_;
  a.key_prev = seed_prev_;
  a.mult = multiplier;
  a.masked = is_remote() ? 1 : 0;
  launch(a, variant, s);
  if (is_remote()) counter_ += range_counter_advance(total, chunk, qr);
}

uint64_t ShareDB::range_counter_advance(size_t total, size_t chunk, size_t rows) {
  uint64_t n = 0;
  for (size_t c0 = 0; c0 < total; c0 += chunk) n += 2 * ceil_div(std::min(chunk, total - c0) * rows * 2, 64);
  return n;
}

GemmArgs ShareDB::packed_args(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                              size_t chunk, uint16_t multiplier, const size_t* rows,
                              uint16_t* const* outs, int n_batches, const uint64_t* ctr_first,
                              int variant, uint64_t* ctr_end) const {
  IRIS_REQUIRE(q.code_length == code_length_ && db.code_length == code_length_, "code length mismatch");
  IRIS_REQUIRE(total > 0 && total % 64 == 0 && chunk % 256 == 0, "bad range geometry");
  IRIS_REQUIRE(offset % 16 == 0 && offset + total <= db.capacity - kMaxBD, "range beyond DB");
  IRIS_REQUIRE(n_batches >= 1 && n_batches <= kMaxSegments, "1 .. 64 batches per packed launch");
  IRIS_REQUIRE(!is_rot_variant(variant) || q.rot.ptr, "rotation-aware variant needs a query uploaded from raw shares");
  GemmArgs a = gemm_args(q, db, variant);
  const GemmTile t = gemm_tile(variant);
  size_t row = 0;
  uint64_t ctr = counter_;
  for (int i = 0; i < n_batches; i++) {
    IRIS_REQUIRE(rows[i] > 0 && outs[i] != nullptr, "empty batch in a packed launch");
    GemmSegment& sg = a.seg[i];
    sg.out = outs[i];
    sg.out_chunk_stride = chunk * rows[i];
    sg.ctr_first = ctr_first ? ctr_first[i] : ctr;
    sg.ctr_per_chunk = 2 * ceil_div(chunk * rows[i] * 2, 64);
    sg.row_begin = (int)row;
    sg.rows = (int)rows[i];
    row += rows[i];
    if (is_remote()) ctr += range_counter_advance(total, chunk, rows[i]);
  }
  IRIS_REQUIRE(row == (size_t)q.rows, "batch rows must add up to the packed query rows");
  a.n_seg = n_batches;
  a.out = outs[0];
  a.n_dtiles = (int)ceil_div(total, (size_t)t.bd);
  a.db_begin = offset;
  a.db_end = offset + total;
  a.chunk = chunk;
  a.nonce = nonce_;
  a.key_own = seed_own_;
  a.key_prev = seed_prev_;
  a.mult = multiplier;
  a.masked = is_remote() ? 1 : 0;
  if (ctr_end) *ctr_end = ctr;
  return a;
}

void ShareDB::dot_reduce_packed(const QuerySet& q, const DbSlice& db, size_t offset, size_t total,
                                size_t chunk, uint16_t multiplier, const size_t* rows,
                                uint16_t* const* outs, int n_batches, const uint64_t* ctr_first,
                                hipStream_t s) {
  DeviceGuard g(device_);
  const int variant = resolve_variant(q);
  uint64_t ctr = 0;
  GemmArgs a = packed_args(q, db, offset, total, chunk, multiplier, rows, outs, n_batches, ctr_first,
                           variant, &ctr);
  launch(a, variant, s);
  if (!ctr_first) counter_ = ctr;
}

namespace {
struct ReshareArg {
  ShareDB* self;
  size_t chunk;
  hipStream_t s;
};
}  // namespace

struct ShareDBTrio {
  static void reshare(ShareDB* const d[3], size_t chunk, hipStream_t const s[3]) {
    // the rows of each party's last dot_reduce (query_length unless a smaller set ran)
    const size_t bytes = chunk * d[0]->last_rows_ * 2;
    for (int p = 0; p < 3; p++) {
      IRIS_REQUIRE(d[p]->peer_id_ == p, "trio members must be parties 0, 1, 2 in order");
      IRIS_REQUIRE(d[p]->last_rows_ == d[0]->last_rows_ && chunk * d[p]->last_rows_ <= d[p]->results_.n,
                   "reshare geometry mismatch");
    }
    const bool copy = d[0]->peer_copy || d[1]->peer_copy || d[2]->peer_copy;
    Event sent[3];
    for (int p = 0; p < 3; p++)
      if (s[(p + 1) % 3] != s[p]) IRIS_HIP_CHECK(hipEventRecord(sent[p].e, s[p]));
    for (int p = 0; p < 3; p++) {
      const int q = (p + 1) % 3;
      // the receiver reads after the sender's GEMM finished (stream order
      // when both run on one stream: a wait there would only cost a marker)
      if (s[q] != s[p]) IRIS_HIP_CHECK(hipStreamWaitEvent(s[q], sent[p].e, 0));
      if (copy) {
        IRIS_HIP_CHECK(hipMemcpyAsync(d[q]->results_peer_.ptr, d[p]->results(), bytes,
                                      hipMemcpyDeviceToDevice, s[q]));
        IRIS_HIP_CHECK(hipEventRecord(d[p]->ev_send_done_.e, s[q]));
        d[q]->peer_view_ = nullptr;
      } else {
        d[q]->peer_view_ = d[p]->results();  // the message is the sender's buffer
      }
    }
    // the sender's stream must not overwrite `results` before its message left
    if (copy)
      for (int p = 0; p < 3; p++) IRIS_HIP_CHECK(hipStreamWaitEvent(s[p], d[p]->ev_send_done_.e, 0));
    for (int p = 0; p < 3; p++)
      if (d[p]->is_remote()) d[p]->counter_ += ceil_div(bytes, 64);  // OTP keystream
  }
};

void ShareDB::reshare_results(size_t chunk, hipStream_t s) {
  IRIS_REQUIRE(ex_ != nullptr, "reshare_results needs a remote engine (exchange handle)");
  ReshareArg arg{this, chunk, s};
  ex_->collective(peer_id_, op_id("reshare", tag_), &arg, [](void* const* args) {
    ShareDB* d[3];
    hipStream_t st[3];
    size_t chunk = static_cast<ReshareArg*>(args[0])->chunk;
    for (int p = 0; p < 3; p++) {
      auto* a = static_cast<ReshareArg*>(args[p]);
      IRIS_REQUIRE(a->chunk == chunk, "parties reshare different chunk sizes");
      d[a->self->peer_id()] = a->self;
      st[a->self->peer_id()] = a->s;
    }
    ShareDB::reshare_trio(d, chunk, st);
  });
}

void ShareDB::reshare_trio(ShareDB* const d[3], size_t chunk, hipStream_t const s[3]) {
  DeviceGuard g(d[0]->device_);
  ShareDBTrio::reshare(d, chunk, s);
}

// ====================================================================== Circuits
Circuits::Circuits(int device, int peer_id, size_t input_size, size_t alloc_size,
                   const ChaChaKey& seed_own, const ChaChaKey& seed_prev, Exchange* ex, int tag)
    : device_(device),
      peer_id_(peer_id),
      tag_(tag),
      alloc_words_(alloc_size),
      chunk_words_(input_size / 64),
      seed_own_(seed_own),
      seed_prev_(seed_prev),
      ex_(ex) {
  IRIS_REQUIRE(peer_id >= 0 && peer_id < 3, "peer_id must be 0..2");
  IRIS_REQUIRE(input_size % 64 == 0, "input size must be a multiple of 64");
  IRIS_REQUIRE(alloc_size >= chunk_words_, "alloc_size must cover the chunk");
  IRIS_REQUIRE(ex != nullptr, "Circuits needs an exchange handle");
  DeviceGuard g(device_);
  const size_t A = alloc_words_ + 8;
  c_a_.alloc(2 * A); c_b_.alloc(2 * A);
  res_a_.alloc(A); res_b_.alloc(A); open_c_.alloc(A);
  corr_a_.alloc(128 * A);
}

void Circuits::set_chunk_size(size_t words) {
  IRIS_REQUIRE(words <= alloc_words_, "chunk size exceeds the allocation");
  chunk_words_ = words;
}

CircuitBufs Circuits::bufs(const uint16_t* ca, const uint16_t* cb, const uint16_t* ma,
                           const uint16_t* mb) const {
  CircuitBufs b{};
  b.code_a = ca; b.code_b = cb; b.mask_a = ma; b.mask_b = mb;
  b.c_a = c_a_.ptr; b.c_b = c_b_.ptr;
  b.corr_a = corr_a_.ptr;
  b.res_a = res_a_.ptr; b.res_b = res_b_.ptr;
  b.lifted_a = lifted_dbg_a; b.lifted_b = lifted_dbg_b;
  b.key_own = seed_own_; b.key_prev = seed_prev_;
  return b;
}

void Circuits::compare_trio(Circuits* const c[3], const uint16_t* const code_a[3],
                            const uint16_t* const code_b[3], const uint16_t* const mask_a[3],
                            const uint16_t* const mask_b[3], size_t len, uint32_t A,
                            hipStream_t const s[3], bool fork) {
  IRIS_REQUIRE(len % 64 == 0, "input length must be a multiple of 64");
  const size_t W = len / 64;
  IRIS_REQUIRE(W % 8 == 0, "elements per call must be a multiple of 512 "
                           "(query_length * chunk / 64 divisible by 8)");
  Exchange* ex = c[0]->ex_;
  DeviceGuard g(ex->device());
  Trio t{};
  for (int p = 0; p < 3; p++) {
    IRIS_REQUIRE(c[p]->peer_id_ == p, "trio members must be parties 0, 1, 2 in order");
    IRIS_REQUIRE(c[p]->ex_ == ex, "trio members must share one exchange");
    IRIS_REQUIRE(W <= c[p]->alloc_words_, "input larger than the Circuits allocation");
    IRIS_REQUIRE(c[p]->counter_ == c[0]->counter_ && c[p]->alloc_words_ == c[0]->alloc_words_,
                 "party RNG counters or allocations diverged");
    t.p[p] = c[p]->bufs(code_a[p], code_b[p], mask_a[p], mask_b[p]);
  }
  t.W = W;
  t.nonce = c[0]->nonce_;
  // The messages between the parties' waves skip the reference's OTP
  // (chacha1 encrypt by the sender, chacha2 decrypt by the receiver,
  // protocol.rs:784-892), which is the identity only when party p's prev seed
  // is party p - 1's own seed; the kernels also rely on it to compute each
  // pairwise ChaCha12 block once.
  for (int p = 0; p < 3; p++)
    IRIS_REQUIRE(std::memcmp(&c[p]->seed_prev_, &c[(p + 2) % 3]->seed_own_, sizeof(ChaChaKey)) == 0,
                 "the three parties' circuit seeds are not pairwise (party p's prev seed must be "
                 "party p - 1's own seed)");
  const CircuitSchedule sc = CircuitSchedule::make(c[0]->counter_, W, c[0]->alloc_words_);
  hipStream_t ps = ex->protocol_stream();
  ex->join_streams(s);
  RippleCtrs lift{};
  lift.csa = sc.lift_csa;
  lift.c[0] = sc.lift_first;
  for (int k = 1; k <= 14; k++) lift.c[k] = sc.lift_round[k];
  lift.c[15] = sc.lift_final;
  LaunchTimer& timer = c[0]->timer;
  if (timer.on()) timer.begin(ps);
  RippleCtrs msb{};
  msb.csa = sc.msb_csa;
  msb.c[0] = sc.msb_first;
  for (int k = 1; k <= 29; k++) msb.c[k] = sc.msb_round[k];
  if (trio_b_is_prev(t)) {
    launch_compare_fused(t, A, lift, sc.inject, msb, ps);
  } else {
    launch_adder_word(t, false, A, lift, ps);
    launch_inject(t, sc.inject, ps);
    launch_adder_word(t, true, A, msb, ps);
  }
  if (timer.on()) timer.end(ps);
  if (fork) ex->fork_streams(s);
  for (int p = 0; p < 3; p++) c[p]->counter_ = sc.end;
}

namespace {
struct CompareArg {
  Circuits* self;
  const uint16_t *ca, *cb, *ma, *mb;
  size_t len;
  uint32_t A;
  hipStream_t s;
  bool fork;
};
struct OpenArg {
  Circuits* self;
  hipStream_t s;
};
}  // namespace

void Circuits::compare_threshold_masked_many(const uint16_t* code_a, const uint16_t* code_b,
                                             const uint16_t* mask_a, const uint16_t* mask_b,
                                             size_t len, uint32_t A, hipStream_t s, bool fork) {
  CompareArg arg{this, code_a, code_b, mask_a, mask_b, len, A, s, fork};
  ex_->collective(peer_id_, op_id("compare_threshold", tag_), &arg, [](void* const* args) {
    Circuits* c[3];
    const uint16_t *ca[3], *cb[3], *ma[3], *mb[3];
    hipStream_t st[3];
    auto* a0 = static_cast<CompareArg*>(args[0]);
    for (int i = 0; i < 3; i++) {
      auto* a = static_cast<CompareArg*>(args[i]);
      IRIS_REQUIRE(a->len == a0->len && a->A == a0->A && a->fork == a0->fork,
                   "parties disagree on length/threshold/fork");
      const int p = a->self->peer_id();
      c[p] = a->self; ca[p] = a->ca; cb[p] = a->cb; ma[p] = a->ma; mb[p] = a->mb; st[p] = a->s;
    }
    compare_trio(c, ca, cb, ma, mb, a0->len, a0->A, st, a0->fork);
  });
}

void Circuits::open_exchange_trio(Circuits* const c[3], hipStream_t const*) {
  // each party's stream reads the previous party's b after the circuits that
  // wrote it (the compare forked the streams after its last launch); the next
  // compare joins the streams before writing it again
  for (int p = 0; p < 3; p++) c[p]->open_view_ = c[(p + 2) % 3]->res_b_.ptr;
}

void Circuits::open_exchange(hipStream_t s) {
  OpenArg arg{this, s};
  ex_->collective(peer_id_, op_id("open", tag_), &arg, [](void* const* args) {
    Circuits* c[3];
    hipStream_t st[3];
    for (int i = 0; i < 3; i++) {
      auto* a = static_cast<OpenArg*>(args[i]);
      c[a->self->peer_id()] = a->self;
      st[a->self->peer_id()] = a->s;
    }
    open_exchange_trio(c, st);
  });
}

// ====================================================================== DistanceComparator
DistanceComparator::DistanceComparator(int device, size_t query_length, size_t max_db_size)
    : device_(device), query_length_(query_length), max_db_size_(max_db_size) {
  IRIS_REQUIRE(query_length % kRotations == 0, "query_length must be batch * 31");
  DeviceGuard g(device_);
  const size_t B = query_length / kRotations;
  match_counters_.alloc(B);
  all_matches_.alloc(kAllMatchesLen * B);
  mcl_.alloc(B);
  mcr_.alloc(B);
  prl_.alloc(kAllMatchesLen * B);
  prr_.alloc(kAllMatchesLen * B);
  cand_.alloc(ceil_div(max_db_size, 64) + 1);
  tiles_.alloc(ceil_div(max_db_size, 64 * 256) + 1);
  ids_.alloc(max_db_size + 1);
  count_.alloc(1);
  cap_counter_.alloc(1);
  cap_q_.alloc(kAllMatchesLen * query_length);
  cap_db_.alloc(kAllMatchesLen * query_length);
  cap_rot_.alloc(kAllMatchesLen * query_length);
}

std::unique_ptr<DevBuf<uint64_t>> DistanceComparator::prepare_db_match_list(size_t db_size) const {
  DeviceGuard g(device_);
  return std::make_unique<DevBuf<uint64_t>>(
      ceil_div(db_size * (query_length_ / kRotations), 64) + 1);
}

void DistanceComparator::set_anon_capture(const AnonCapture& c, const uint64_t* host_targets,
                                          size_t n_targets) {
  DeviceGuard g(device_);
  if (!c.enabled()) {
    anon_ = AnonCapture{};
    return;
  }
  IRIS_REQUIRE(c.code_a && c.code_b && c.mask_a && c.mask_b && c.idx, "capture buffers are NULL");
  IRIS_REQUIRE(c.max_query_length % kRotations == 0 && c.max_query_length > 0,
               "max_query_length must be a multiple of 31");
  IRIS_REQUIRE(!n_targets || host_targets, "targets are NULL");
  anon_ = c;
  if (n_targets) {
    if (anon_targets_.n < n_targets) anon_targets_.alloc(n_targets, false);
    IRIS_HIP_CHECK(hipMemcpy(anon_targets_.ptr, host_targets, n_targets * 8, hipMemcpyHostToDevice));
  }
  anon_.targets = anon_targets_.ptr;
  anon_.n_targets = n_targets;
}

void DistanceComparator::open_results(const uint64_t* a, const uint64_t* b, const uint64_t* c,
                                      uint64_t* bitmap, size_t chunk_len, size_t real_chunk_len,
                                      size_t offset, size_t total_db_len, size_t batch_size,
                                      hipStream_t s, size_t rows, const uint32_t* slot_of,
                                      const DotViews* dots) const {
  DeviceGuard g(device_);
  if (rows == 0) rows = query_length_;
  IRIS_REQUIRE(chunk_len % 64 == 0, "chunk length must be a multiple of 64");
  IRIS_REQUIRE(batch_size * kRotations <= rows && rows <= query_length_, "batch larger than the query rows");
  OpenArgs o{};
  o.a = a; o.b = b; o.c = c; o.bitmap = bitmap; o.slot_of = slot_of;
  o.n_words = chunk_len * rows / 64;
  o.chunk_len = chunk_len;
  o.query_len = batch_size * kRotations;
  o.offset = offset;
  o.real_chunk_len = real_chunk_len;
  o.total_db_len = total_db_len;
  const bool capture = dots && anon_.enabled();
  if (capture) {
    o.anon = anon_;
    o.dots = *dots;
  }
  launch_open(o, s);
  if (capture) {
    AnonTargetArgs t{};
    t.cap = anon_; t.dots = *dots; t.a = a; t.b = b; t.c = c;
    t.chunk_len = chunk_len; t.real_chunk_len = real_chunk_len; t.offset = offset;
    t.batch = batch_size; t.index_map = nullptr;
    launch_anon_targets(t, s);
  }
}

size_t DistanceComparator::get_partial_results(const uint64_t* bitmap, size_t db_size,
                                               size_t batch_size, const uint32_t* extra_host,
                                               size_t n_extra, hipStream_t s, size_t row_bits) {
  DeviceGuard g(device_);
  IRIS_REQUIRE(db_size <= max_db_size_, "db_size beyond the comparator's max_db_size");
  IRIS_REQUIRE(batch_size * kRotations <= query_length_, "batch larger than the query rows");
  if (row_bits == 0) row_bits = db_size;
  IRIS_REQUIRE(row_bits >= db_size, "bitmap rows shorter than the DB");
  if (n_extra) {
    if (extra_.n < n_extra) extra_.alloc(n_extra, false);
    IRIS_HIP_CHECK(hipMemcpyAsync(extra_.ptr, extra_host, n_extra * 4, hipMemcpyHostToDevice, s));
  }
  CandidateArgs c{};
  c.bitmap = bitmap; c.row_bits = row_bits; c.batch = batch_size; c.db_size = db_size;
  c.extra = extra_.ptr; c.n_extra = n_extra;
  c.cand = cand_.ptr; c.tiles = tiles_.ptr; c.ids = ids_.ptr; c.count = count_.ptr;
  launch_candidates(c, s);
  uint32_t n = 0;
  IRIS_HIP_CHECK(hipMemcpyAsync(&n, count_.ptr, 4, hipMemcpyDeviceToHost, s));
  IRIS_HIP_CHECK(hipStreamSynchronize(s));
  return n;
}

void DistanceComparator::open_results_mapped(const uint64_t* a, const uint64_t* b,
                                             const uint64_t* c, uint64_t* bitmap,
                                             size_t chunk_len, size_t real_chunk_len,
                                             size_t total_db_len, size_t batch_size,
                                             const uint32_t* index_map, hipStream_t s,
                                             size_t rows, const uint32_t* slot_of,
                                             const DotViews* dots) {
  DeviceGuard g(device_);
  if (rows == 0) rows = query_length_;
  IRIS_REQUIRE(chunk_len % 64 == 0, "chunk length must be a multiple of 64");
  IRIS_REQUIRE(real_chunk_len <= chunk_len, "real chunk longer than the chunk");
  IRIS_REQUIRE(batch_size * kRotations <= rows && rows <= query_length_, "batch larger than the query rows");
  MappedOpenArgs o{};
  o.a = a; o.b = b; o.c = c; o.bitmap = bitmap;
  o.n_words = chunk_len * rows / 64;
  o.chunk_len = chunk_len;
  o.query_len = batch_size * kRotations;
  o.real_chunk_len = real_chunk_len;
  o.total_db_len = total_db_len;
  o.index_map = index_map;
  o.slot_of = slot_of;
  o.counter = cap_counter_.ptr;
  o.cap_query = cap_q_.ptr;
  o.cap_db = cap_db_.ptr;
  o.cap_rot = cap_rot_.ptr;
  o.capacity = cap_q_.n;
  o.set_bits = bitmap != nullptr;
  const bool capture = dots && anon_.enabled();
  IRIS_REQUIRE(bitmap || capture, "nothing to open into: no bitmap and no distance capture");
  if (capture) {
    o.anon = anon_;
    o.dots = *dots;
  }
  launch_open_mapped(o, s);
  if (capture) {
    AnonTargetArgs t{};
    t.cap = anon_; t.dots = *dots; t.a = a; t.b = b; t.c = c;
    t.chunk_len = chunk_len; t.real_chunk_len = real_chunk_len; t.offset = 0;
    t.batch = batch_size; t.index_map = index_map;
    launch_anon_targets(t, s);
  }
}

size_t DistanceComparator::fetch_partial_rotations(uint32_t* query, uint32_t* db, int8_t* rot,
                                                   size_t capacity, hipStream_t s) {
  DeviceGuard g(device_);
  uint32_t n = 0;
  IRIS_HIP_CHECK(hipMemcpyAsync(&n, cap_counter_.ptr, 4, hipMemcpyDeviceToHost, s));
  IRIS_HIP_CHECK(hipStreamSynchronize(s));
  const size_t m = std::min<size_t>({(size_t)n, capacity, cap_q_.n});
  if (m) {
    IRIS_HIP_CHECK(hipMemcpyAsync(query, cap_q_.ptr, m * 4, hipMemcpyDeviceToHost, s));
    IRIS_HIP_CHECK(hipMemcpyAsync(db, cap_db_.ptr, m * 4, hipMemcpyDeviceToHost, s));
    IRIS_HIP_CHECK(hipMemcpyAsync(rot, cap_rot_.ptr, m, hipMemcpyDeviceToHost, s));
  }
  fill_u32(cap_counter_.ptr, 1, 0, s);
  IRIS_HIP_CHECK(hipStreamSynchronize(s));
  return n;
}

void DistanceComparator::join_db_matches(const uint64_t* left, const uint64_t* right,
                                         uint32_t* final_results, size_t db_size,
                                         const uint64_t* or_policy, size_t max_db_size,
                                         hipStream_t s) {
  DeviceGuard g(device_);
  MergeArgs m{};
  m.left = left; m.right = right; m.or_policy = or_policy; m.final_results = final_results;
  m.match_counters = match_counters_.ptr; m.all_matches = all_matches_.ptr;
  m.mcl = mcl_.ptr; m.mcr = mcr_.ptr; m.prl = prl_.ptr; m.prr = prr_.ptr;
  m.query_len = query_length_ / kRotations;
  m.db_len = db_size;
  m.max_db_len = max_db_size;
  launch_merge(m, s);
}

void DistanceComparator::join_batch_matches(const uint64_t* left, const uint64_t* right,
                                            uint32_t* final_results, size_t batch_size,
                                            size_t row_bits, hipStream_t s) {
  DeviceGuard g(device_);
  if (row_bits == 0) row_bits = batch_size;
  IRIS_REQUIRE(batch_size * kRotations <= query_length_, "batch larger than the query rows");
  IRIS_REQUIRE(row_bits >= batch_size, "batch bitmap rows shorter than the batch");
  BatchMergeArgs m{};
  m.left = left; m.right = right; m.final_results = final_results;
  m.match_counters = match_counters_.ptr; m.all_matches = all_matches_.ptr;
  m.batch = batch_size; m.row_bits = row_bits;
  launch_merge_batch(m, s);
}

void DistanceComparator::reset(hipStream_t s) {
  DeviceGuard g(device_);
  const size_t B = query_length_ / kRotations;
  fill_u32(match_counters_.ptr, B, 0, s);
  fill_u32(mcl_.ptr, B, 0, s);
  fill_u32(mcr_.ptr, B, 0, s);
}

}  // namespace irismpc
