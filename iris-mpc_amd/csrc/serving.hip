// Request loops of the serving and per-batch forms (host code over the
// engines; the device work is the scan step of actor.hip).
#include "serving.hpp"

#include <algorithm>
#include <cstring>
#include <limits>
#include <set>

#include "comparator.hpp"
#include "host_crypto.hpp"
#include "synth.hpp"

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

// sharing randomness of the default iris (code 0, mask all ones)
constexpr uint64_t kDummyDataset = 0x0de1e7e;
// device memory a request adds to the active set: limbs and staging of its
// 31 query rows, code and mask, three parties
constexpr size_t kRequestFootprint = 3 * (size_t)kRotations * (kCodeLen + kMaskLen) * 2 * 2;
constexpr size_t kMaxQueuedSteps = 2;

ChaChaKey derive(const ChaChaKey& k, uint64_t nonce) {
  ChaChaKey o;
  derive_seed(k.k, nonce, o.k);
  return o;
}

// Adds the seconds between construction and destruction to *acc (under mu).
struct PhaseTimer {
  double* acc;
  std::mutex* mu;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  PhaseTimer(double* a, std::mutex* m) : acc(a), mu(m) {}
  ~PhaseTimer() {
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::lock_guard<std::mutex> lk(*mu);
    *acc += dt;
  }
};

bool event_done(hipEvent_t e) {
  const hipError_t r = hipEventQuery(e);
  if (r == hipSuccess) return true;
  if (r == hipErrorNotReady) return false;
  IRIS_HIP_CHECK(r);
  return false;
}
}  // namespace

// ====================================================================== EngineTrio
EngineTrio::EngineTrio(int device, size_t maxq, size_t db_capacity, const ChaChaKey own[3],
                       const ChaChaKey prev[3], int tag0, uint64_t nonce_base, size_t chunk_rows)
    : max_queries(maxq) {
  DeviceGuard g(device);
  ex = std::make_unique<Exchange>(device, 300.0);
  const size_t rows = maxq * kRotations;
  const std::vector<uint16_t> zc(maxq * kCodeLen, 0), zm(maxq * kMaskLen, 0);
  for (int p = 0; p < 3; p++) {
    IRIS_HIP_CHECK(hipStreamCreateWithFlags(&s[p], hipStreamNonBlocking));
    codes[p] = std::make_unique<ShareDB>(device, p, chunk_rows, rows, kCodeLen, derive(own[p], nonce_base),
                                         derive(prev[p], nonce_base), ex.get(), tag0);
    masks[p] = std::make_unique<ShareDB>(device, p, chunk_rows, rows, kMaskLen, derive(own[p], nonce_base + 1),
                                         derive(prev[p], nonce_base + 1), ex.get(), tag0 + 1);
    circ[p] = std::make_unique<Circuits>(device, p, rows * chunk_rows, rows * chunk_rows / 64,
                                         derive(own[p], nonce_base + 5), derive(prev[p], nonce_base + 5),
                                         ex.get(), tag0 + 2);
    cmp[p] = std::make_unique<DistanceComparator>(device, rows, db_capacity);
    qc[p] = codes[p]->upload_query_shares_rows(zc.data(), maxq, rows, s[p]);
    qm[p] = masks[p]->upload_query_shares_rows(zm.data(), maxq, rows, s[p]);
  }
  sync();
}

EngineTrio::~EngineTrio() {
  for (auto st : s)
    if (st) (void)hipStreamSynchronize(st);
  for (int p = 0; p < 3; p++) {
    qc[p].reset();
    qm[p].reset();
    cmp[p].reset();
    circ[p].reset();
    masks[p].reset();
    codes[p].reset();
  }
  ex.reset();
  for (auto st : s)
    if (st) (void)hipStreamDestroy(st);
}

void EngineTrio::sync() const {
  for (auto st : s) IRIS_HIP_CHECK(hipStreamSynchronize(st));
}

void EngineTrio::parties(DbSlice* const code_dbs[3], DbSlice* const mask_dbs[3], PartyScan out[3]) const {
  for (int p = 0; p < 3; p++)
    out[p] = PartyScan{codes[p].get(), masks[p].get(), circ[p].get(), cmp[p].get(),
                       code_dbs[p],    mask_dbs[p],    qc[p].get(),   qm[p].get()};
}

// ====================================================================== Serving
Serving::Serving(int device, const ChaChaKey party_seeds[3], DbSlice* const code_dbs[3],
                 DbSlice* const mask_dbs[3], size_t n_rows, size_t capacity, size_t max_active,
                 uint32_t threshold_a)
    : device_(device),
      capacity_(capacity),
      max_active_(max_active),
      words_(ceil_div(std::max<size_t>(capacity, 1), 64)),
      threshold_a_(threshold_a),
      n_rows_(n_rows) {
  IRIS_REQUIRE(max_active >= 1 && max_active <= 64, "1 .. 64 active queries");
  active_cap_ = max_active;
  IRIS_REQUIRE(capacity > (size_t)kDbChunk, "DB capacity must exceed one chunk");
  IRIS_REQUIRE(n_rows <= capacity - kDbChunk, "DB rows beyond the capacity");
  for (int p = 0; p < 3; p++) {
    IRIS_REQUIRE(code_dbs[p] && mask_dbs[p], "three code and three mask DB slices");
    IRIS_REQUIRE(code_dbs[p]->code_length == kCodeLen && mask_dbs[p]->code_length == kMaskLen,
                 "code DB slices of 12,800 and mask DB slices of 6,400 coefficients");
    IRIS_REQUIRE(code_dbs[p]->capacity >= capacity + kMaxBD && mask_dbs[p]->capacity >= capacity + kMaxBD,
                 "DB slices smaller than the capacity");
    code_dbs_[p] = code_dbs[p];
    mask_dbs_[p] = mask_dbs[p];
  }
  DeviceGuard g(device);
  ChaChaKey own[3], prev[3];
  for (int p = 0; p < 3; p++) {
    own[p] = party_seeds[p];
    prev[p] = party_seeds[(p + 2) % 3];
  }
  // Each engine set derives its code, mask and circuit keys at nonces base,
  // base + 1 and base + 5: bases more than 5 apart give the sets disjoint
  // ChaCha12 keystreams. A zero-share mask repeated across two computations
  // would let the receiver of both masked values learn their difference.
  constexpr uint64_t kScanNonce = 0, kReauthNonce = 16, kSelfNonce = 32;
  static_assert(kReauthNonce > kScanNonce + 5 && kSelfNonce > kReauthNonce + 5,
                "two engine sets would share a keystream");
  scan_ = std::make_unique<EngineTrio>(device, max_active, capacity, own, prev, 0, kScanNonce);
  // reauth: one query over a 64-row window, its own keystreams
  small_ = std::make_unique<EngineTrio>(device, 1, capacity, own, prev, 8, kReauthNonce);
  // the self-comparison of requests in flight: up to max_active queries
  // against the set's irises (one step of a 512-row DB), its own keystreams
  self_ = std::make_unique<EngineTrio>(device, max_active, kSelfChunk, own, prev, 24, kSelfNonce, kSelfChunk);
  scan_->parties(code_dbs_, mask_dbs_, scan_p_);
  small_->parties(code_dbs_, mask_dbs_, small_p_);
  {
    DbSlice *sc[3], *sm[3];
    for (int p = 0; p < 3; p++) {
      self_code_db_[p] = self_->codes[p]->alloc_db(kSelfChunk);
      self_mask_db_[p] = self_->masks[p]->alloc_db(kSelfChunk);
      sc[p] = self_code_db_[p].get();
      sm[p] = self_mask_db_[p].get();
      self_raw_code_[p].alloc(max_active * kCodeLen, false);
      self_raw_mask_[p].alloc(max_active * kMaskLen, false);
    }
    self_->parties(sc, sm, self_p_);
  }
  for (int p = 0; p < 3; p++) {
    bitmaps_[p].alloc(max_active * words_ + 2);
    small_bitmaps_[p].alloc(words_ + 2);
    self_bitmaps_[p].alloc(max_active + 2);
    slot_maps_[p].alloc(max_active);
    dummy_code_[p].alloc(kCodeLen);
    dummy_mask_[p].alloc(kMaskLen);
  }
  {  // shares of the default iris
    DevBuf<uint64_t> code(kIrisWords), mask(kIrisWords);
    hipStream_t s = scan_->s[0];
    fill_u64(mask.ptr, kIrisWords, ~0ull, s);
    uint16_t* co[3] = {dummy_code_[0].ptr, dummy_code_[1].ptr, dummy_code_[2].ptr};
    uint16_t* mo[3] = {dummy_mask_[0].ptr, dummy_mask_[1].ptr, dummy_mask_[2].ptr};
    launch_synth_shares(code.ptr, mask.ptr, 1, kDummyDataset, 0, co, mo, s);
    dummy_code_host_.resize(3 * kCodeLen);
    dummy_mask_host_.resize(3 * kMaskLen);
    for (int p = 0; p < 3; p++) {
      IRIS_HIP_CHECK(hipMemcpyAsync(dummy_code_host_.data() + p * kCodeLen, co[p], kCodeLen * 2,
                                    hipMemcpyDeviceToHost, s));
      IRIS_HIP_CHECK(hipMemcpyAsync(dummy_mask_host_.data() + p * kMaskLen, mo[p], kMaskLen * 2,
                                    hipMemcpyDeviceToHost, s));
    }
    IRIS_HIP_CHECK(hipStreamSynchronize(s));
  }
  for (int k = 0; k < kRefillRing; k++) {
    refill_code_[k].alloc(3 * max_active * kCodeLen);
    refill_mask_[k].alloc(3 * max_active * kMaskLen);
  }
  for (int k = 0; k < kSlotRing; k++) slot_host_[k].alloc(max_active);
  for (int k = 0; k < kWriteRing; k++) {
    write_code_[k].alloc(3 * kCodeLen);
    write_mask_[k].alloc(3 * kMaskLen);
    write_dev_code_[k].alloc(3 * kCodeLen, false);
    write_dev_mask_[k].alloc(3 * kMaskLen, false);
  }
  readout_.alloc(max_active * words_);
  for (int k = 0; k < kSelfRing; k++) self_host_[k].alloc(max_active);
  small_code_.alloc(3 * kCodeLen);
  small_mask_.alloc(3 * kMaskLen);
  for (int p = 0; p < 3; p++) {
    IRIS_HIP_CHECK(hipEventCreateWithFlags(&last_write_[p], hipEventDisableTiming));
    IRIS_HIP_CHECK(hipEventRecord(last_write_[p], scan_->s[p]));
  }
  scan_->sync();
  t0_ = std::chrono::steady_clock::now();
}

Serving::~Serving() {
  try {
    if (running_) finish(60.0);
  } catch (...) {
  }
  {  // requests left unfinished by a stop that failed
    std::lock_guard<std::mutex> lk(mu_);
    draining_ = reauth_stop_ = true;
  }
  cv_.notify_all();
  reauth_cv_.notify_all();
  if (loop_.joinable()) loop_.join();
  if (reauth_.joinable()) reauth_.join();
  if (scan_) scan_->sync();
  if (small_) small_->sync();
  for (auto& st : inflight_)
    for (auto e : st.e) (void)hipEventDestroy(e);
  for (auto e : recycle_later_) (void)hipEventDestroy(e);
  for (int k = 0; k < kRefillRing; k++)
    for (auto e : refill_ev_[k]) (void)hipEventDestroy(e);
  for (int k = 0; k < kSlotRing; k++)
    for (auto e : slot_ev_[k]) (void)hipEventDestroy(e);
  for (int k = 0; k < kWriteRing; k++)
    for (auto e : write_ev_[k]) (void)hipEventDestroy(e);
  for (auto& kv : reqs_)
    if (kv.second->readout) (void)hipEventDestroy(kv.second->readout);
  for (auto e : event_pool_) (void)hipEventDestroy(e);
  for (auto e : last_write_)
    if (e) (void)hipEventDestroy(e);
}

double Serving::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

size_t Serving::n_rows() const {
  std::lock_guard<std::mutex> lk(mu_);
  return n_rows_;
}

ServingStats Serving::stats() const {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

hipEvent_t Serving::event() {
  if (!event_pool_.empty()) {
    hipEvent_t e = event_pool_.back();
    event_pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  IRIS_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return e;
}

void Serving::recycle(hipEvent_t e) {
  if (e) event_pool_.push_back(e);
}

// ---------------------------------------------------------------- requests
uint64_t Serving::submit(int kind, const uint16_t* code, const uint16_t* mask, int64_t target,
                         double arrival) {
  IRIS_REQUIRE(kind >= kReqUniqueness && kind <= kReqDelete, "unknown request kind");
  auto r = std::make_unique<Req>();
  r->kind = kind;
  r->target = target;
  if (kind != kReqDelete) {
    IRIS_REQUIRE(code && mask, "request without shares");
    r->code.assign(code, code + 3 * kCodeLen);
    r->mask.assign(mask, mask + 3 * kMaskLen);
  }
  r->res.arrival = arrival < 0 ? now() : arrival;
  std::lock_guard<std::mutex> lk(mu_);
  IRIS_REQUIRE(!loop_error_, "the request loop has failed");
  IRIS_REQUIRE(!draining_, "the request loop is stopping");
  const uint64_t id = next_id_++;
  r->id = id;
  incoming_.push_back(r.get());
  reqs_[id] = std::move(r);
  n_submitted_++;
  cv_.notify_one();
  return id;
}

bool Serving::result(uint64_t id, Result* out) const {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) return false;
  const Result& r = it->second->res;
  out->state = r.state;
  out->arrival = r.arrival;
  if (r.state != kReqPending) *out = r;  // written before complete() published the state
  return true;
}

// A finished request's shares are dropped at once (they were copied to the
// staging buffers before its last step); its result stays until released.
void Serving::complete(Req* r, int state, const std::string& err) {
  std::lock_guard<std::mutex> lk(mu_);
  if (r->res.state != kReqPending) return;
  r->res.done = now();
  if (!err.empty()) r->res.error = err;
  r->res.state = state;
  std::vector<uint16_t>().swap(r->code);
  std::vector<uint16_t>().swap(r->mask);
  n_finished_++;
  done_cv_.notify_all();
}

bool Serving::release(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  IRIS_REQUIRE(it != reqs_.end(), "unknown request id");
  if (it->second->res.state == kReqPending) return false;
  if (it->second->readout) recycle_later_.push_back(it->second->readout);
  reqs_.erase(it);
  return true;
}

void Serving::fail_all(const std::string& msg) {
  std::lock_guard<std::mutex> lk(mu_);
  const double t = now();
  for (auto& kv : reqs_) {
    Result& res = kv.second->res;
    if (res.state != kReqPending) continue;
    res.state = kReqFailed;
    res.error = msg;
    res.done = t;
    n_finished_++;  // the shares stay: a queue may still hold the request until the loop ends
  }
  done_cv_.notify_all();
}

// Moves the submitted requests whose arrival time has come to *out (in
// submission order); with `wait` and none due, first waits up to 2 ms or
// until the next arrival. Returns whether the loop is stopping.
bool Serving::take_incoming(std::vector<Req*>* out, bool wait) {
  std::unique_lock<std::mutex> lk(mu_);
  const size_t n0 = out->size();
  auto release = [&] {
    const double t = now();
    size_t kept = 0;
    for (Req* r : incoming_) {
      if (r->res.arrival <= t)
        out->push_back(r);
      else
        incoming_[kept++] = r;
    }
    incoming_.resize(kept);
  };
  release();
  if (wait && out->size() == n0 && !draining_) {
    double next = std::numeric_limits<double>::infinity();
    for (const Req* r : incoming_) next = std::min(next, r->res.arrival);
    const double dt = std::min(0.002, std::max(0.0, next - now()));
    cv_.wait_for(lk, std::chrono::duration<double>(dt));
    release();
  }
  return draining_;
}

void Serving::start(int policy, size_t max_batch, size_t reserve_bytes) {
  IRIS_REQUIRE(policy == kPolicyContinuous || policy == kPolicyPerBatch, "unknown serving policy");
  IRIS_REQUIRE(policy == kPolicyContinuous || (max_batch >= 1 && max_batch <= max_active_),
               "per-batch form: 1 .. max_active queries per batch");
  IRIS_REQUIRE(!running_, "the request loop is running");
  {
    std::lock_guard<std::mutex> lk(mu_);
    IRIS_REQUIRE(!loop_error_, "the request loop has failed");
    draining_ = reauth_stop_ = false;
    stats_ = ServingStats{};
  }
  running_ = true;
  reauth_ = std::thread([this] { reauth_main(); });
  loop_ = std::thread([this, policy, max_batch, reserve_bytes] { loop_main(policy, max_batch, reserve_bytes); });
}

void Serving::finish(double timeout_s) {
  IRIS_REQUIRE(running_, "the request loop is not running");
  bool timed_out = false;
  {
    std::unique_lock<std::mutex> lk(mu_);
    auto all_done = [&] { return n_finished_ == n_submitted_; };
    if (timeout_s > 0)
      timed_out = !done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), all_done);
    else
      done_cv_.wait(lk, all_done);
    draining_ = reauth_stop_ = true;
    if (timed_out) {
      abort_ = true;
      if (!loop_error_)
        loop_error_ = std::make_exception_ptr(Error("requests still pending when the wait timed out"));
    }
  }
  cv_.notify_all();
  reauth_cv_.notify_all();
  loop_.join();
  reauth_.join();
  running_ = false;
  if (timed_out) fail_all("the request loop stopped before the request finished");
  std::exception_ptr err;
  {
    std::lock_guard<std::mutex> lk(mu_);
    err = loop_error_;
  }
  if (err) std::rethrow_exception(err);
}

void Serving::loop_main(int policy, size_t max_batch, size_t reserve_bytes) {
  try {
    IRIS_HIP_CHECK(hipSetDevice(device_));
    if (policy == kPolicyContinuous)
      loop_continuous(reserve_bytes);
    else
      loop_per_batch(max_batch);
    drain_steps();
    for (auto& r : self_rounds_) recycle(r.done);  // a loop left early: rounds nobody reads
    self_rounds_.clear();
  } catch (const std::exception& e) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      loop_error_ = std::current_exception();
    }
    fail_all(std::string("request loop failed: ") + e.what());
  }
}

// Validates and applies an insertion or deletion: the rows are written on the
// parties' scan streams (after the queued steps, before the later ones); the
// reauth worker's reads wait for them through last_write_. A request that
// fails validation is completed with its error; device errors propagate.
void Serving::apply_write(Req* r) {
  size_t row, rows;
  {
    std::lock_guard<std::mutex> lk(mu_);
    rows = n_rows_;
  }
  if (r->kind == kReqInsert) {
    row = rows;
    if (row >= capacity_ - kDbChunk) {
      r->res.db_rows = rows;
      complete(r, kReqFailed,
               "DB capacity exhausted (" + std::to_string(row) + " rows of " + std::to_string(capacity_) + ")");
      return;
    }
  } else {
    if (r->target < 0 || (size_t)r->target >= rows) {
      r->res.db_rows = rows;
      complete(r, kReqFailed,
               "deletion of row " + std::to_string(r->target) + " outside the DB (" + std::to_string(rows) +
                   " rows)");
      return;
    }
    row = (size_t)r->target;
  }
  {
    std::lock_guard<std::mutex> lk(db_mu_);
    if (r->kind == kReqInsert) {
      const int k = write_next_;
      write_next_ = (write_next_ + 1) % kWriteRing;
      for (auto e : write_ev_[k]) {
        IRIS_HIP_CHECK(hipEventSynchronize(e));
        recycle(e);
      }
      write_ev_[k].clear();
      std::memcpy(write_code_[k].ptr, r->code.data(), 3 * kCodeLen * 2);
      std::memcpy(write_mask_[k].ptr, r->mask.data(), 3 * kMaskLen * 2);
      for (int p = 0; p < 3; p++) {
        hipStream_t s = scan_->s[p];
        uint16_t* dc = write_dev_code_[k].ptr + p * kCodeLen;
        uint16_t* dm = write_dev_mask_[k].ptr + p * kMaskLen;
        IRIS_HIP_CHECK(hipMemcpyAsync(dc, write_code_[k].ptr + p * kCodeLen, kCodeLen * 2,
                                      hipMemcpyHostToDevice, s));
        IRIS_HIP_CHECK(hipMemcpyAsync(dm, write_mask_[k].ptr + p * kMaskLen, kMaskLen * 2,
                                      hipMemcpyHostToDevice, s));
        scan_->codes[p]->load_records_device(*code_dbs_[p], row, 1, dc, s);
        scan_->masks[p]->load_records_device(*mask_dbs_[p], row, 1, dm, s);
        hipEvent_t e = event();
        IRIS_HIP_CHECK(hipEventRecord(e, s));
        write_ev_[k].push_back(e);
        IRIS_HIP_CHECK(hipEventRecord(last_write_[p], s));
      }
    } else {
      for (int p = 0; p < 3; p++) {
        hipStream_t s = scan_->s[p];
        scan_->codes[p]->load_records_device(*code_dbs_[p], row, 1, dummy_code_[p].ptr, s);
        scan_->masks[p]->load_records_device(*mask_dbs_[p], row, 1, dummy_mask_[p].ptr, s);
        IRIS_HIP_CHECK(hipEventRecord(last_write_[p], s));
      }
    }
    // under db_mu_ as well: a reauth reads the version with the write it sees
    std::lock_guard<std::mutex> lk2(mu_);
    if (r->kind == kReqInsert) n_rows_++;
    r->res.db_rows = n_rows_;
    r->res.db_version = ++n_writes_;
  }
  r->res.value = (int64_t)row;
  complete(r, kReqDone);
}

void Serving::clear_chunk(int slot, size_t chunk) {
  const size_t w0 = (size_t)slot * words_ + chunk * (kDbChunk / 64);
  const size_t n = std::min<size_t>(kDbChunk / 64, words_ - chunk * (kDbChunk / 64));
  for (int p = 0; p < 3; p++) fill_u64(bitmaps_[p].ptr + w0, n, 0, scan_->s[p]);
}

void Serving::patch_step(const std::vector<Req*>& active, size_t chunk, size_t rows) {
  for (Req* a : active) {
    if (chunk >= a->need) {  // an insertion opened a chunk past the request's
      a->covered.resize(chunk + 1, 0);
      a->need = chunk + 1;
    } else if (a->covered[chunk]) {
      // ordered after the queued steps that set these bits
      clear_chunk(a->slot, chunk);
      a->covered[chunk] = 0;
      a->n_covered--;
    }
  }
  step(active.size(), chunk, 1, rows);
  for (Req* a : active)
    if (!a->covered[chunk]) {
      a->covered[chunk] = 1;
      a->n_covered++;
    }
}

// ---------------------------------------------------------------- reauth
void Serving::reauth_main() {
  try {
    IRIS_HIP_CHECK(hipSetDevice(device_));
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> lk(mu_);
    if (!loop_error_) loop_error_ = std::current_exception();
    return;
  }
  for (;;) {
    Req* r;
    {
      std::unique_lock<std::mutex> lk(mu_);
      reauth_cv_.wait(lk, [&] { return !reauth_q_.empty() || reauth_stop_; });
      if (reauth_q_.empty() || abort_) return;
      r = reauth_q_.front();
      reauth_q_.pop_front();
    }
    try {
      serve_reauth(r);
    } catch (const std::exception& e) {
      complete(r, kReqFailed, e.what());
    }
  }
}

// One query x 31 rotations against DB row `target` on the one-query engine
// set: the chunk step over the 64-row window around the target.
void Serving::serve_reauth(Req* r) {
  const size_t rows = n_rows();
  r->res.db_rows = rows;
  if (r->target < 0 || (size_t)r->target >= rows) {
    complete(r, kReqFailed,
             "reauth target " + std::to_string(r->target) + " outside the DB (" + std::to_string(rows) +
                 " rows)");
    return;
  }
  const size_t target = (size_t)r->target, off = target / 64 * 64;
  std::memcpy(small_code_.ptr, r->code.data(), 3 * kCodeLen * 2);
  std::memcpy(small_mask_.ptr, r->mask.data(), 3 * kMaskLen * 2);
  for (int p = 0; p < 3; p++) {
    hipStream_t s = small_->s[p];
    small_->codes[p]->refill_query_shares_async(*small_->qc[p], small_code_.ptr + p * kCodeLen, 1, kRotations, s,
                                                false);
    small_->masks[p]->refill_query_shares_async(*small_->qm[p], small_mask_.ptr + p * kMaskLen, 1, kRotations, s,
                                                false);
    fill_u64(small_bitmaps_[p].ptr, words_ + 2, 0, s);
  }
  uint64_t word = 0;
  {
    std::lock_guard<std::mutex> lk(db_mu_);  // the window's rows are not written meanwhile
    {
      std::lock_guard<std::mutex> lk2(mu_);
      r->res.db_version = n_writes_;
    }
    for (int p = 0; p < 3; p++) IRIS_HIP_CHECK(hipStreamWaitEvent(small_->s[p], last_write_[p], 0));
    uint64_t* const bms[3] = {small_bitmaps_[0].ptr, small_bitmaps_[1].ptr, small_bitmaps_[2].ptr};
    scan_chunk_trio(small_p_, off, off + 64, 1, threshold_a_, bms, nullptr, 0, small_->s);
    IRIS_HIP_CHECK(hipMemcpyAsync(&word, small_bitmaps_[0].ptr + target / 64, 8, hipMemcpyDeviceToHost,
                                  small_->s[0]));
    small_->sync();
  }
  r->res.value = (int64_t)((word >> (target % 64)) & 1);
  complete(r, kReqDone);
}

// ---------------------------------------------------------------- device steps
int Serving::refill_scan(const std::vector<const Req*>& reqs, size_t batch, size_t rows) {
  IRIS_REQUIRE(batch >= 1 && batch <= max_active_ && reqs.size() <= batch, "query set beyond the slots");
  PhaseTimer pt(&stats_.t_refill, &mu_);
  {
    std::lock_guard<std::mutex> lk(mu_);
    stats_.refills++;
  }
  const int k = refill_next_;
  refill_next_ = (refill_next_ + 1) % kRefillRing;
  for (auto e : refill_ev_[k]) {
    IRIS_HIP_CHECK(hipEventSynchronize(e));
    recycle(e);
  }
  refill_ev_[k].clear();
  for (int p = 0; p < 3; p++) {
    uint16_t* bc = refill_code_[k].ptr + (size_t)p * max_active_ * kCodeLen;
    uint16_t* bm = refill_mask_[k].ptr + (size_t)p * max_active_ * kMaskLen;
    for (size_t j = 0; j < batch; j++) {
      const Req* r = j < reqs.size() ? reqs[j] : nullptr;
      const uint16_t* c = r ? r->code.data() : dummy_code_host_.data();
      const uint16_t* m = r ? r->mask.data() : dummy_mask_host_.data();
      std::memcpy(bc + j * kCodeLen, c + p * kCodeLen, kCodeLen * 2);
      std::memcpy(bm + j * kMaskLen, m + p * kMaskLen, kMaskLen * 2);
    }
    hipStream_t s = scan_->s[p];
    // the scan reads only the query rows: no DB form of the active set
    scan_->codes[p]->refill_query_shares_async(*scan_->qc[p], bc, batch, rows, s, false);
    scan_->masks[p]->refill_query_shares_async(*scan_->qm[p], bm, batch, rows, s, false);
    hipEvent_t e = event();
    IRIS_HIP_CHECK(hipEventRecord(e, s));
    refill_ev_[k].push_back(e);
  }
  return k;
}

void Serving::self_compare(int k, const std::vector<uint64_t>& ids, size_t q0,
                           const std::vector<std::pair<Req*, size_t>>& fresh) {
  const size_t n = ids.size(), nq = n - q0;
  IRIS_REQUIRE(q0 >= 1 && q0 < n && n <= max_active_, "self-comparison: queries after the first of 2 .. max_active");
  const int buf = (int)(self_seq_ % kSelfRing);
  // the round that last used this readout buffer is decoded first
  if (self_rounds_.size() >= (size_t)kSelfRing) collect_self(self_rounds_.front().seq);
  for (int p = 0; p < 3; p++) {
    hipStream_t s = self_->s[p];
    const uint16_t* bc = refill_code_[k].ptr + (size_t)p * max_active_ * kCodeLen;
    const uint16_t* bm = refill_mask_[k].ptr + (size_t)p * max_active_ * kMaskLen;
    // the set's irises (positions 0 .. n) as the DB, the queries at q0 ..
    IRIS_HIP_CHECK(hipMemcpyAsync(self_raw_code_[p].ptr, bc, n * kCodeLen * 2, hipMemcpyHostToDevice, s));
    IRIS_HIP_CHECK(hipMemcpyAsync(self_raw_mask_[p].ptr, bm, n * kMaskLen * 2, hipMemcpyHostToDevice, s));
    self_->codes[p]->load_records_device(*self_code_db_[p], 0, n, self_raw_code_[p].ptr, s);
    self_->masks[p]->load_records_device(*self_mask_db_[p], 0, n, self_raw_mask_[p].ptr, s);
    self_->codes[p]->refill_query_shares_async(*self_->qc[p], bc + q0 * kCodeLen, nq, nq * kRotations, s, false);
    self_->masks[p]->refill_query_shares_async(*self_->qm[p], bm + q0 * kMaskLen, nq, nq * kRotations, s, false);
    hipEvent_t e = event();  // the staging ring entry is reused after these reads
    IRIS_HIP_CHECK(hipEventRecord(e, s));
    refill_ev_[k].push_back(e);
    fill_u64(self_bitmaps_[p].ptr, nq + 2, 0, s);
  }
  uint64_t* const bms[3] = {self_bitmaps_[0].ptr, self_bitmaps_[1].ptr, self_bitmaps_[2].ptr};
  // bit j of row i: query q0 + i (31 rotations) against iris j
  scan_chunk_trio(self_p_, 0, n, nq, threshold_a_, bms, nullptr, 64, self_->s);
  hipStream_t s0 = self_->s[0];
  IRIS_HIP_CHECK(hipMemcpyAsync(self_host_[buf].ptr, self_bitmaps_[0].ptr, nq * 8, hipMemcpyDeviceToHost, s0));
  SelfRound r;
  r.seq = ++self_seq_;
  r.buf = buf;
  r.done = event();
  IRIS_HIP_CHECK(hipEventRecord(r.done, s0));
  r.ids = ids;
  r.q0 = q0;
  r.fresh = fresh;
  for (auto& f : fresh) f.first->self_round = r.seq;
  self_rounds_.push_back(std::move(r));
}

void Serving::collect_self(uint64_t upto) {
  while (!self_rounds_.empty()) {
    SelfRound& r = self_rounds_.front();
    if (r.seq > upto && !event_done(r.done)) break;
    IRIS_HIP_CHECK(hipEventSynchronize(r.done));
    const uint64_t* rows = self_host_[r.buf].ptr;
    for (const auto& f : r.fresh) {
      Req* q = f.first;
      const size_t b = f.second;
      // above the diagonal only: the requests admitted before it
      for (size_t j = 0; j < b; j++)
        if (r.ids[j] && ((rows[b - r.q0] >> j) & 1)) q->res.batch_matches.push_back(r.ids[j]);
      q->self_round = 0;
    }
    recycle(r.done);
    self_rounds_.pop_front();
  }
}

void Serving::set_slots(const std::vector<int>& slots) {
  if (slots == slots_) return;
  const int k = slot_next_;
  slot_next_ = (slot_next_ + 1) % kSlotRing;
  for (auto e : slot_ev_[k]) {
    IRIS_HIP_CHECK(hipEventSynchronize(e));
    recycle(e);
  }
  slot_ev_[k].clear();
  for (size_t j = 0; j < slots.size(); j++) slot_host_[k].ptr[j] = (uint32_t)slots[j];
  for (int p = 0; p < 3; p++) {
    hipStream_t s = scan_->s[p];
    IRIS_HIP_CHECK(hipMemcpyAsync(slot_maps_[p].ptr, slot_host_[k].ptr, slots.size() * 4,
                                  hipMemcpyHostToDevice, s));
    hipEvent_t e = event();
    IRIS_HIP_CHECK(hipEventRecord(e, s));
    slot_ev_[k].push_back(e);
  }
  slots_ = slots;
}

void Serving::clear_slot(int slot) {
  for (int p = 0; p < 3; p++)
    fill_u64(bitmaps_[p].ptr + (size_t)slot * words_, words_, 0, scan_->s[p]);
}

void Serving::step(size_t batch, size_t chunk_index, size_t n_chunks, size_t db_rows) {
  IRIS_REQUIRE(batch >= 1 && batch * n_chunks <= max_active_ && n_chunks <= kMaxStepChunks,
               "step beyond the result buffers");
  {
    PhaseTimer pt(&stats_.t_wait, &mu_);
    while (inflight_.size() >= kMaxQueuedSteps) {
      for (auto e : inflight_.front().e) {
        IRIS_HIP_CHECK(hipEventSynchronize(e));
        recycle(e);
      }
      inflight_.pop_front();
    }
  }
  PhaseTimer pt(&stats_.t_enqueue, &mu_);
  uint64_t* const bms[3] = {bitmaps_[0].ptr, bitmaps_[1].ptr, bitmaps_[2].ptr};
  const uint32_t* const so[3] = {slot_maps_[0].ptr, slot_maps_[1].ptr, slot_maps_[2].ptr};
  scan_chunks_trio(scan_p_, chunk_index * kDbChunk, n_chunks, db_rows, batch, threshold_a_, bms, so,
                   words_ * 64, scan_->s);
  StepEvents ev;
  for (int p = 0; p < 3; p++) {
    ev.e[p] = event();
    IRIS_HIP_CHECK(hipEventRecord(ev.e[p], scan_->s[p]));
  }
  inflight_.push_back(ev);
  std::lock_guard<std::mutex> lk(mu_);
  stats_.steps++;
}

void Serving::drain_steps() {
  scan_->sync();
  for (auto& st : inflight_)
    for (auto e : st.e) recycle(e);
  inflight_.clear();
}

hipEvent_t Serving::begin_readout(int slot) {
  hipStream_t s = scan_->s[0];
  IRIS_HIP_CHECK(hipMemcpyAsync(readout_.ptr + (size_t)slot * words_, bitmaps_[0].ptr + (size_t)slot * words_,
                                words_ * 8, hipMemcpyDeviceToHost, s));
  hipEvent_t e = event();
  IRIS_HIP_CHECK(hipEventRecord(e, s));
  return e;
}

std::vector<uint32_t> Serving::readout_rows(int slot, size_t db_rows) const {
  std::vector<uint32_t> rows;
  const uint64_t* w = readout_.ptr + (size_t)slot * words_;
  const size_t n = ceil_div(db_rows, 64);
  for (size_t i = 0; i < n; i++) {
    uint64_t x = w[i];
    while (x) {
      const size_t row = i * 64 + (size_t)__builtin_ctzll(x);
      if (row < db_rows) rows.push_back((uint32_t)row);
      x &= x - 1;
    }
  }
  return rows;
}

// ---------------------------------------------------------------- step sizes
void Serving::set_admission(size_t active_cap, size_t queue_threshold) {
  IRIS_REQUIRE(!running_, "admission changes only while the request loop is stopped");
  IRIS_REQUIRE(active_cap >= 1, "an active cap of at least one request");
  active_cap_ = std::min(active_cap, max_active_);
  queue_threshold_ = queue_threshold;
}

void Serving::set_step_table(const uint32_t* buckets, const uint32_t* chunks, size_t n) {
  IRIS_REQUIRE(!running_, "step sizes change only while the request loop is stopped");
  std::vector<std::pair<uint32_t, uint32_t>> t;
  for (size_t i = 0; i < n; i++) {
    IRIS_REQUIRE(buckets[i] >= 1 && chunks[i] >= 1 && chunks[i] <= kMaxStepChunks, "step table entry out of range");
    t.emplace_back(buckets[i], chunks[i]);
  }
  std::sort(t.begin(), t.end());
  step_buckets_.clear();
  step_chunks_.clear();
  for (auto& e : t) {
    step_buckets_.push_back(e.first);
    step_chunks_.push_back(e.second);
  }
}

// Chunks one step covers: several while few queries are active (their
// launches alone would leave the GPU partly idle), one under load so that
// arrivals join at short intervals; bounded by the result buffers (chunks x
// rows within the engines' query rows) and the ring's end. With a table, the
// entry of the largest bucket at or below `batch` (the smallest bucket's
// below all); otherwise 8 / 4 / 2 / 1 chunks for 1 / 2 / <= 8 / more queries.
size_t Serving::chunks_per_step(size_t batch, size_t chunk_index, size_t n_chunks) const {
  size_t k;
  if (!step_buckets_.empty()) {
    size_t i = 0;
    for (size_t j = 0; j < step_buckets_.size(); j++)
      if (step_buckets_[j] <= batch) i = j;
    k = step_chunks_[i];
  } else {
    k = batch == 1 ? 8 : batch <= 2 ? 4 : batch <= 8 ? 2 : 1;
  }
  k = std::min({k, std::max<size_t>(1, max_active_ / std::max<size_t>(batch, 1)),
                n_chunks > chunk_index ? n_chunks - chunk_index : (size_t)1, (size_t)kMaxStepChunks});
  return std::max<size_t>(1, k);
}

double Serving::time_step(size_t batch, size_t n_chunks, int reps) {
  IRIS_REQUIRE(!running_, "time_step needs the request loop stopped");
  IRIS_REQUIRE(batch >= 1 && batch * n_chunks <= max_active_ && n_chunks >= 1 && n_chunks <= kMaxStepChunks,
               "step beyond the result buffers");
  DeviceGuard g(device_);
  refill_scan({}, batch, batch * kRotations);
  std::vector<int> slots(batch);
  for (size_t j = 0; j < batch; j++) slots[j] = (int)j;
  set_slots(slots);
  double best = std::numeric_limits<double>::infinity();
  for (int r = 0; r <= std::max(reps, 1); r++) {
    drain_steps();
    const auto t0 = std::chrono::steady_clock::now();
    step(batch, 0, n_chunks, n_rows_);
    drain_steps();
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (r > 0) best = std::min(best, t);
  }
  return best;
}

std::vector<uint32_t> Serving::run_alone(const uint16_t* code, const uint16_t* mask) {
  IRIS_REQUIRE(!running_, "run_alone needs the request loop stopped");
  DeviceGuard g(device_);
  Req q;
  q.code.assign(code, code + 3 * kCodeLen);
  q.mask.assign(mask, mask + 3 * kMaskLen);
  refill_scan({&q}, 1, kRotations);
  set_slots({0});
  clear_slot(0);
  const size_t rows = n_rows_, n = n_chunks_of(rows);
  for (size_t c = 0; c < n;) {
    const size_t k = chunks_per_step(1, c, n);
    step(1, c, k, rows);
    c += k;
  }
  hipEvent_t e = begin_readout(0);
  IRIS_HIP_CHECK(hipEventSynchronize(e));
  recycle(e);
  drain_steps();
  return readout_rows(0, rows);
}

// ---------------------------------------------------------------- loops
void Serving::loop_continuous(size_t reserve_bytes) {
  std::deque<Req*> waiting, reading;
  std::vector<Req*> active, in;
  std::vector<int> free_slots;
  for (size_t i = 0; i < max_active_; i++) free_slots.push_back((int)i);
  std::vector<uint64_t> active_ids;
  std::set<size_t> written;  // chunks written since the last step
  size_t cursor = 0;
  for (;;) {
    // finished readouts, oldest first
    while (!reading.empty() && event_done(reading.front()->readout)) {
      Req* r = reading.front();
      reading.pop_front();
      {
        PhaseTimer pt(&stats_.t_decode, &mu_);
        r->res.matches = readout_rows(r->slot, r->res.db_rows);
      }
      r->res.value = (int64_t)r->res.matches.size();
      recycle(r->readout);
      r->readout = nullptr;
      free_slots.push_back(r->slot);
      if (r->self_round) collect_self(r->self_round);
      complete(r, kReqDone);
    }
    collect_self(0);  // the finished ones
    const bool idle = active.empty() && waiting.empty() && reading.empty();
    in.clear();
    const bool stopping = take_incoming(&in, idle);
    for (Req* r : in) {
      if (r->kind == kReqUniqueness) {
        waiting.push_back(r);
      } else if (r->kind == kReqReauth) {
        std::lock_guard<std::mutex> lk(mu_);
        reauth_q_.push_back(r);
        reauth_cv_.notify_one();
      } else {
        apply_write(r);
        if (r->res.state == kReqDone) written.insert((size_t)r->res.value / kDbChunk);
      }
    }
    if ((stopping && in.empty() && idle) || abort_) break;
    // admission between chunk steps: up to active_cap requests, or every
    // waiting one while more than queue_threshold wait (slots and HBM allowing)
    std::vector<std::pair<Req*, size_t>> fresh;  // admitted now, with their positions
    while (!waiting.empty() && !free_slots.empty() &&
           (active.size() < active_cap_ || waiting.size() > queue_threshold_)) {
      if (!active.empty()) {  // with nothing active one request always joins
        size_t free = 0, total = 0;
        {
          PhaseTimer pt(&stats_.t_admit, &mu_);
          IRIS_HIP_CHECK(hipMemGetInfo(&free, &total));
        }
        if (free < reserve_bytes + kRequestFootprint * (active.size() + 1)) {
          std::lock_guard<std::mutex> lk(mu_);
          stats_.admission_waits++;
          break;
        }
      }
      Req* r = waiting.front();
      waiting.pop_front();
      r->slot = free_slots.front();
      free_slots.erase(free_slots.begin());
      r->res.db_rows = n_rows();
      r->need = n_chunks_of(r->res.db_rows);
      r->covered.assign(r->need, 0);
      r->n_covered = 0;
      clear_slot(r->slot);
      if (!active.empty()) fresh.emplace_back(r, active.size());
      active.push_back(r);
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      stats_.max_concurrent = std::max<uint64_t>(stats_.max_concurrent, active.size());
    }
    if (active.empty()) {
      written.clear();  // later requests scan the DB as it is
      if (!reading.empty() && waiting.empty()) IRIS_HIP_CHECK(hipEventSynchronize(reading.front()->readout));
      continue;
    }
    std::vector<uint64_t> ids;
    std::vector<int> slots;
    for (Req* r : active) {
      ids.push_back(r->id);
      slots.push_back(r->slot);
    }
    if (ids != active_ids) {
      std::vector<const Req*> q(active.begin(), active.end());
      const int k = refill_scan(q, active.size(), active.size() * kRotations);
      active_ids = ids;
      // the newly admitted requests (the set's last positions, in admission
      // order) against every request admitted before them and still scanning
      if (!fresh.empty()) self_compare(k, ids, fresh.front().second, fresh);
    }
    set_slots(slots);
    const size_t rows = n_rows(), n = n_chunks_of(rows);
    // every active request answers for the DB as it stands at its last step:
    // a written chunk is scanned again for the whole active set right away
    for (size_t c : written) patch_step(active, c, rows);
    written.clear();
    cursor %= n;
    const size_t k = chunks_per_step(active.size(), cursor, n);
    step(active.size(), cursor, k, rows);
    for (Req* r : active)
      for (size_t c = cursor; c < cursor + k; c++)
        if (c < r->need && !r->covered[c]) {
          r->covered[c] = 1;
          r->n_covered++;
        }
    cursor += k;
    for (size_t i = 0; i < active.size();) {
      Req* r = active[i];
      if (r->n_covered == r->need) {  // its bitmap row is copied out after its last step
        r->res.db_rows = rows;
        {
          std::lock_guard<std::mutex> lk(mu_);
          r->res.db_version = n_writes_;
        }
        r->readout = begin_readout(r->slot);
        reading.push_back(r);
        active.erase(active.begin() + i);
      } else {
        i++;
      }
    }
  }
}

// The reference's batch after batch (actor.rs:584-644): the uniqueness and
// reauth requests that arrived while a batch ran form the next batch (at most
// max_batch); its query set is padded to max_batch x 31 rows with the default
// iris, as the reference sizes its queries and GEMMs by max_batch_size
// whatever the batch holds (actor.rs:895-901, share_db.rs:800-845); every
// batch is one scan of the whole DB, one chunk per step. A reauth is answered
// by its row of the batch's scan at its target. Insertions and deletions that
// arrived before a batch starts are applied first.
void Serving::loop_per_batch(size_t max_batch) {
  std::deque<Req*> pending;
  std::vector<Req*> in;
  for (;;) {
    in.clear();
    const bool stopping = take_incoming(&in, pending.empty());
    for (Req* r : in) pending.push_back(r);
    if (abort_) break;
    if (pending.empty()) {
      if (stopping) break;
      continue;
    }
    std::vector<Req*> batch;
    while (!pending.empty() && batch.size() < max_batch) {
      Req* r = pending.front();
      pending.pop_front();
      if (r->kind == kReqInsert || r->kind == kReqDelete)
        apply_write(r);
      else
        batch.push_back(r);
    }
    if (batch.empty()) continue;
    std::vector<const Req*> q(batch.begin(), batch.end());
    // the batch's uniqueness requests against the earlier ones of the batch
    std::vector<uint64_t> uids(batch.size(), 0);
    std::vector<std::pair<Req*, size_t>> fresh;
    for (size_t j = 0; j < batch.size(); j++)
      if (batch[j]->kind == kReqUniqueness) {
        uids[j] = batch[j]->id;
        if (j > 0) fresh.emplace_back(batch[j], j);
      }
    const int k = refill_scan(q, max_batch, max_batch * kRotations);
    if (!fresh.empty()) self_compare(k, uids, fresh.front().second, fresh);
    const size_t rows = n_rows(), n = n_chunks_of(rows);
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (Req* r : batch) r->res.db_version = n_writes_;
    }
    std::vector<int> slots(batch.size());
    for (size_t j = 0; j < batch.size(); j++) slots[j] = (int)j;
    set_slots(slots);
    for (int s : slots) clear_slot(s);
    for (size_t c = 0; c < n; c++) step(batch.size(), c, 1, rows);
    {
      std::lock_guard<std::mutex> lk(mu_);
      stats_.batches++;
    }
    std::vector<hipEvent_t> evs;
    for (int s : slots) evs.push_back(begin_readout(s));
    for (size_t j = 0; j < batch.size(); j++) {
      IRIS_HIP_CHECK(hipEventSynchronize(evs[j]));
      recycle(evs[j]);
      Req* r = batch[j];
      r->res.db_rows = rows;
      std::vector<uint32_t> m = readout_rows(slots[j], rows);
      if (r->kind == kReqReauth) {
        if (r->target < 0 || (size_t)r->target >= rows) {
          complete(r, kReqFailed,
                   "reauth target " + std::to_string(r->target) + " outside the DB (" + std::to_string(rows) +
                       " rows)");
          continue;
        }
        r->res.value = std::binary_search(m.begin(), m.end(), (uint32_t)r->target) ? 1 : 0;
      } else {
        r->res.value = (int64_t)m.size();
        r->res.matches = std::move(m);
        if (r->self_round) collect_self(r->self_round);
      }
      complete(r, kReqDone);
    }
  }
}

}  // namespace irismpc
