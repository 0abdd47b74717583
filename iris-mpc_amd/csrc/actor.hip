// Per-party DB scan (chunk loop) and its three-party driver.
#include "actor.hpp"

#include <exception>
#include <thread>
#include <mutex>
#include <functional>
#include <condition_variable>

namespace irismpc {

static DotViews dot_views(const PartyScan& p) {
  DotViews d;
  d.code_a = p.codes->results();
  d.code_b = p.codes->results_peer();
  d.mask_a = p.masks->results();
  d.mask_b = p.masks->results_peer();
  return d;
}

// Several consecutive chunks in one step: each chunk's GEMMs write its own
// region of the result buffer ([rows][chunk width], side by side), one reshare
// and one compare cover them all, and each chunk opens from its own result
// words. The opened bits are those of one step per chunk (the circuit masks
// differ). With a small query set this fills the GPU with fewer, larger
// launches.
void scan_chunks(const PartyScan& p, size_t offset, size_t n_chunks, size_t db_size,
                 size_t batch_size, uint32_t threshold_a, uint64_t* bitmap,
                 const uint32_t* slot_of, size_t row_bits, hipStream_t s) {
  const size_t rows = (size_t)p.code_query->rows;
  IRIS_REQUIRE((size_t)p.mask_query->rows == rows, "code/mask query sets disagree on rows");
  IRIS_REQUIRE(batch_size * kRotations <= rows, "batch larger than the query rows");
  IRIS_REQUIRE(n_chunks >= 1 && n_chunks <= kMaxStepChunks, "1 .. 16 chunks per step");
  // every opened bit lands at offset + c * chunk + d < db_size within its row:
  // rows shorter than the DB would spill into the next query's row
  IRIS_REQUIRE(row_bits == 0 || row_bits >= db_size, "bitmap rows shorter than the DB");
  // actor.rs:2331-2337 pads to 64; the circuits need a multiple of 8 words
  const size_t pad = rows % 8 == 0 ? 64 : 512;
  size_t real[kMaxStepChunks], width[kMaxStepChunks], start[kMaxStepChunks];
  size_t total = 0, n = 0;
  for (size_t c = 0; c < n_chunks; c++) {
    const size_t off = offset + c * kDbChunk;
    // actor.rs:2286-2291, 2318-2533: the chunk covers rows [off, off + real)
    real[c] = db_size > off ? std::min<size_t>(kDbChunk, db_size - off) : 0;
    if (c > 0 && real[c] == 0) break;  // nothing past the DB's end
    width[c] = round_up(std::max<size_t>(real[c], 1), pad);
    start[c] = total;
    total += width[c];
    n = c + 1;
  }
  // every chunk but the last is kDbChunk wide, so one launch per matrix covers
  // them with the per-chunk layout and masks
  p.codes->dot_reduce_chunks(*p.code_query, *p.code_db, offset, total, kDbChunk, 1, s);
  p.masks->dot_reduce_chunks(*p.mask_query, *p.mask_db, offset, total, kDbChunk, 2, s);
  // the distance capture reads the peer shares on this party's stream after
  // the circuits: they need a copy the peer's next GEMM cannot overwrite
  p.codes->peer_copy = p.masks->peer_copy = p.comparator->anon_capture_enabled();
  p.codes->reshare_results(total, s);
  p.masks->reshare_results(total, s);
  const size_t len = total * rows;
  p.circuits->set_chunk_size(len / 64);
  p.circuits->compare_threshold_masked_many(p.codes->results(), p.codes->results_peer(),
                                            p.masks->results(), p.masks->results_peer(), len,
                                            threshold_a, s);
  p.circuits->open_exchange(s);
  for (size_t c = 0; c < n; c++) {
    if (real[c] == 0) continue;
    const size_t e0 = start[c] * rows, w0 = e0 / 64;
    DotViews dots = dot_views(p);
    dots.code_a += e0;
    dots.code_b += e0;
    dots.mask_a += e0;
    dots.mask_b += e0;
    p.comparator->open_results(p.circuits->result_a() + w0, p.circuits->result_b() + w0,
                               p.circuits->result_c() + w0, bitmap, width[c], real[c],
                               offset + c * kDbChunk, row_bits ? row_bits : db_size, batch_size, s,
                               rows, slot_of, &dots);
  }
}

void scan_chunk(const PartyScan& p, size_t offset, size_t db_size, size_t batch_size,
                uint32_t threshold_a, uint64_t* bitmap, const uint32_t* slot_of, size_t row_bits,
                hipStream_t s) {
  scan_chunks(p, offset, 1, db_size, batch_size, threshold_a, bitmap, slot_of, row_bits, s);
}

void scan_subset(const PartyScan& p, DbSlice& code_sub, DbSlice& mask_sub, const uint32_t* ids,
                 size_t n_ids, size_t total_db_len, size_t batch_size, uint32_t threshold_a,
                 uint64_t* bitmap, hipStream_t s, uint32_t anon_threshold) {
  const size_t rows = (size_t)p.code_query->rows;
  IRIS_REQUIRE((size_t)p.mask_query->rows == rows, "code/mask query sets disagree on rows");
  IRIS_REQUIRE(batch_size * kRotations <= rows, "batch larger than the query rows");
  const size_t pad = rows % 8 == 0 ? 64 : 512;
  IRIS_REQUIRE(round_up(std::min<size_t>(n_ids, kDbChunk), pad) <=
                   std::min(code_sub.capacity, mask_sub.capacity),
               "subset slices hold fewer than DB_CHUNK_SIZE rows");
  // actor.rs:1955-1975: subset chunks of DB_CHUNK_SIZE ids
  for (size_t c0 = 0; c0 < n_ids; c0 += kDbChunk) {
    const size_t real = std::min<size_t>(kDbChunk, n_ids - c0);
    const size_t dot_chunk = round_up(real, pad);  // actor.rs:1977-1980
    p.codes->prefetch_subset(*p.code_db, ids + c0, real, code_sub, s);
    p.masks->prefetch_subset(*p.mask_db, ids + c0, real, mask_sub, s);
    p.codes->dot_reduce(*p.code_query, code_sub, dot_chunk, 0, 1, s);
    p.masks->dot_reduce(*p.mask_query, mask_sub, dot_chunk, 0, 2, s);
    p.codes->peer_copy = p.masks->peer_copy = p.comparator->anon_capture_enabled();
    p.codes->reshare_results(dot_chunk, s);
    p.masks->reshare_results(dot_chunk, s);
    const size_t len = dot_chunk * rows;
    p.circuits->set_chunk_size(len / 64);
    if (anon_threshold && p.comparator->anon_capture_enabled()) {  // actor.rs:2040-2120
      const DotViews dots = dot_views(p);
      p.circuits->compare_threshold_masked_many(p.codes->results(), p.codes->results_peer(),
                                                p.masks->results(), p.masks->results_peer(), len,
                                                anon_threshold, s);
      p.circuits->open_exchange(s);
      p.comparator->open_results_mapped(p.circuits->result_a(), p.circuits->result_b(),
                                        p.circuits->result_c(), nullptr, dot_chunk, real,
                                        total_db_len, batch_size, ids + c0, s, rows, nullptr,
                                        &dots);
    }
    p.circuits->compare_threshold_masked_many(p.codes->results(), p.codes->results_peer(),
                                              p.masks->results(), p.masks->results_peer(), len,
                                              threshold_a, s);
    p.circuits->open_exchange(s);
    p.comparator->open_results_mapped(p.circuits->result_a(), p.circuits->result_b(),
                                      p.circuits->result_c(), bitmap, dot_chunk, real,
                                      total_db_len, batch_size, ids + c0, s, rows);
  }
}

void scan_batch_self(const PartyScan& p, size_t batch_size, uint32_t threshold_a,
                     uint64_t* bitmap, size_t row_bits, hipStream_t s) {
  IRIS_REQUIRE(p.code_query->insert && p.mask_query->insert,
               "the batch self-comparison needs query sets uploaded from raw shares");
  IRIS_REQUIRE(batch_size >= 1 && batch_size <= (size_t)p.code_query->batch &&
                   batch_size <= (size_t)p.mask_query->batch,
               "batch larger than the uploaded queries");
  PartyScan b = p;
  b.code_db = p.code_query->insert.get();
  b.mask_db = p.mask_query->insert.get();
  scan_chunk(b, 0, batch_size, batch_size, threshold_a, bitmap, nullptr,
             row_bits ? row_bits : batch_size, s);
}

// The chunk loop with two result buffers, as the reference alternates two
// streams (actor.rs:2318-2533): chunk c's GEMMs run on the party stream `s`
// into result slot c % 2 while chunk c - 1's circuits and open run on the
// exchange's protocol stream. A slot is reused once the open of the chunk
// that last read it has run (event on the protocol stream); the circuits of a
// chunk wait for the three parties' GEMMs (the compare joins the streams).
static void scan_db_pipelined(const PartyScan& p, size_t db_size, size_t batch_size,
                              uint32_t threshold_a, uint64_t* bitmap, hipStream_t s) {
  const size_t rows = (size_t)p.code_query->rows;
  IRIS_REQUIRE((size_t)p.mask_query->rows == rows, "code/mask query sets disagree on rows");
  IRIS_REQUIRE(batch_size * kRotations <= rows, "batch larger than the query rows");
  hipStream_t ps = p.circuits->exchange()->protocol_stream();
  Event freed[2];
  bool used[2] = {false, false};
  const size_t pad = rows % 8 == 0 ? 64 : 512;
  size_t offset = 0;
  int c = 0;
  p.codes->peer_copy = p.masks->peer_copy = false;
  do {
    const int slot = c & 1;
    const size_t real = db_size > offset ? std::min<size_t>(kDbChunk, db_size - offset) : 0;
    const size_t dot_chunk = round_up(std::max<size_t>(real, 1), pad);
    if (used[slot]) IRIS_HIP_CHECK(hipStreamWaitEvent(s, freed[slot].e, 0));
    p.codes->set_slot(slot);
    p.masks->set_slot(slot);
    p.codes->dot_reduce(*p.code_query, *p.code_db, dot_chunk, offset, 1, s);
    p.masks->dot_reduce(*p.mask_query, *p.mask_db, dot_chunk, offset, 2, s);
    p.codes->reshare_results(dot_chunk, s);
    p.masks->reshare_results(dot_chunk, s);
    const size_t len = dot_chunk * rows;
    p.circuits->set_chunk_size(len / 64);
    p.circuits->compare_threshold_masked_many(p.codes->results(), p.codes->results_peer(),
                                              p.masks->results(), p.masks->results_peer(), len,
                                              threshold_a, s, /*fork=*/false);
    p.circuits->open_exchange(s);
    if (real > 0)
      p.comparator->open_results(p.circuits->result_a(), p.circuits->result_b(),
                                 p.circuits->result_c(), bitmap, dot_chunk, real, offset, db_size,
                                 batch_size, ps, rows);
    IRIS_HIP_CHECK(hipEventRecord(freed[slot].e, ps));
    used[slot] = true;
    offset += kDbChunk;
    c++;
  } while (offset < db_size);
  // the caller's stream covers the whole scan
  Event done;
  IRIS_HIP_CHECK(hipEventRecord(done.e, ps));
  IRIS_HIP_CHECK(hipStreamWaitEvent(s, done.e, 0));
  p.codes->set_slot(0);
  p.masks->set_slot(0);
}

void scan_db(const PartyScan& p, size_t db_size, size_t batch_size, uint32_t threshold_a,
             uint64_t* bitmap, hipStream_t s) {
  if (!p.comparator->anon_capture_enabled()) {
    scan_db_pipelined(p, db_size, batch_size, threshold_a, bitmap, s);
    return;
  }
  size_t offset = 0;
  do {
    scan_chunk(p, offset, db_size, batch_size, threshold_a, bitmap, nullptr, 0, s);
    offset += kDbChunk;
  } while (offset < db_size);
}

// Three party threads kept per calling thread: a trio call wakes them instead
// of creating threads (creating three threads and their HIP state costs about
// 0.1 ms, a tenth of a one-query chunk step). Each calling thread has its own
// three, so a request loop and a reauth worker never share them.
class TrioThreads {
 public:
  TrioThreads() {
    for (int i = 0; i < 3; i++) th_[i] = std::thread([this, i] { loop(i); });
  }
  ~TrioThreads() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++gen_;
    }
    go_.notify_all();
    for (auto& t : th_) t.join();
  }
  // Runs body(0), body(1), body(2) on the three threads; rethrows the first
  // error after all three returned.
  void run(const std::function<void(int)>& body, int device) {
    std::unique_lock<std::mutex> lk(mu_);
    body_ = &body;
    device_ = device;
    pending_ = 3;
    for (auto& e : err_) e = nullptr;
    ++gen_;
    go_.notify_all();
    done_.wait(lk, [&] { return pending_ == 0; });
    body_ = nullptr;
    for (auto& e : err_)
      if (e) std::rethrow_exception(e);
  }

 private:
  void loop(int i) {
    uint64_t seen = 0;
    for (;;) {
      const std::function<void(int)>* body;
      int device;
      {
        std::unique_lock<std::mutex> lk(mu_);
        go_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
        body = body_;
        device = device_;
      }
      try {
        IRIS_HIP_CHECK(hipSetDevice(device));
        (*body)(i);
      } catch (...) {
        err_[i] = std::current_exception();
      }
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::thread th_[3];
  std::mutex mu_;
  std::condition_variable go_, done_;
  const std::function<void(int)>* body_ = nullptr;
  int device_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::exception_ptr err_[3];
};

template <class F>
static void run_trio(F&& body) {
  int dev = 0;
  IRIS_HIP_CHECK(hipGetDevice(&dev));
  thread_local TrioThreads threads;
  const std::function<void(int)> fn = [&](int i) { body(i); };
  threads.run(fn, dev);
}

void scan_chunk_trio(const PartyScan p[3], size_t offset, size_t db_size, size_t batch_size,
                     uint32_t threshold_a, uint64_t* const bitmaps[3],
                     const uint32_t* const slot_of[3], size_t row_bits, hipStream_t const s[3]) {
  run_trio([&](int i) {
    scan_chunk(p[i], offset, db_size, batch_size, threshold_a, bitmaps[i],
               slot_of ? slot_of[i] : nullptr, row_bits, s[i]);
  });
}

void scan_chunks_trio(const PartyScan p[3], size_t offset, size_t n_chunks, size_t db_size,
                      size_t batch_size, uint32_t threshold_a, uint64_t* const bitmaps[3],
                      const uint32_t* const slot_of[3], size_t row_bits, hipStream_t const s[3]) {
  run_trio([&](int i) {
    scan_chunks(p[i], offset, n_chunks, db_size, batch_size, threshold_a, bitmaps[i],
                slot_of ? slot_of[i] : nullptr, row_bits, s[i]);
  });
}

void scan_subset_trio(const PartyScan p[3], DbSlice* const code_sub[3], DbSlice* const mask_sub[3],
                      const uint32_t* const ids[3], size_t n_ids, size_t total_db_len,
                      size_t batch_size, uint32_t threshold_a, uint64_t* const bitmaps[3],
                      hipStream_t const s[3], uint32_t anon_threshold) {
  run_trio([&](int i) {
    scan_subset(p[i], *code_sub[i], *mask_sub[i], ids[i], n_ids, total_db_len, batch_size,
                threshold_a, bitmaps[i], s[i], anon_threshold);
  });
}

void scan_batch_self_trio(const PartyScan p[3], size_t batch_size, uint32_t threshold_a,
                          uint64_t* const bitmaps[3], size_t row_bits, hipStream_t const s[3]) {
  run_trio([&](int i) { scan_batch_self(p[i], batch_size, threshold_a, bitmaps[i], row_bits, s[i]); });
}

// The three parties' chunk loop from one host thread. Per chunk the six share
// GEMMs (three parties x code, mask) run as one launch on s[0] into result
// slot c % 2 (one kernel boundary per chunk instead of six);
// the reshares hand over views, and the circuits and opens of chunk c run on
// the exchange's protocol stream while chunk c + 1's GEMMs run (a slot is
// reused once the open of the chunk that last read it has run). Shares,
// counters and bitmaps are those of scan_db called by the three parties.
static void scan_db_trio_ordered(const PartyScan p[3], size_t db_size, size_t batch_size,
                                 uint32_t threshold_a, uint64_t* const bitmaps[3],
                                 hipStream_t const s[3]) {
  const size_t rows = (size_t)p[0].code_query->rows;
  Exchange* ex = p[0].circuits->exchange();
  for (int i = 0; i < 3; i++) {
    IRIS_REQUIRE(p[i].codes->peer_id() == i && p[i].masks->peer_id() == i && p[i].circuits->peer_id() == i,
                 "trio members must be parties 0, 1, 2 in order");
    IRIS_REQUIRE((size_t)p[i].code_query->rows == rows && (size_t)p[i].mask_query->rows == rows,
                 "the parties' query sets disagree on rows");
    IRIS_REQUIRE(p[i].circuits->exchange() == ex, "trio members must share one exchange");
  }
  IRIS_REQUIRE(batch_size * kRotations <= rows, "batch larger than the query rows");
  hipStream_t g = s[0], ps = ex->protocol_stream();
  const hipStream_t gs[3] = {g, g, g};
  {  // the GEMM stream starts after the callers' earlier work on s[1], s[2]
    Event e1, e2;
    IRIS_HIP_CHECK(hipEventRecord(e1.e, s[1]));
    IRIS_HIP_CHECK(hipEventRecord(e2.e, s[2]));
    IRIS_HIP_CHECK(hipStreamWaitEvent(g, e1.e, 0));
    IRIS_HIP_CHECK(hipStreamWaitEvent(g, e2.e, 0));
  }
  ShareDB *cd[3], *md[3];
  Circuits* cc[3];
  for (int i = 0; i < 3; i++) {
    cd[i] = p[i].codes;
    md[i] = p[i].masks;
    cc[i] = p[i].circuits;
    cd[i]->peer_copy = md[i]->peer_copy = false;
  }
  Event freed[2];
  bool used[2] = {false, false};
  const size_t pad = rows % 8 == 0 ? 64 : 512;
  size_t offset = 0;
  int c = 0;
  do {
    const int slot = c & 1;
    const size_t real = db_size > offset ? std::min<size_t>(kDbChunk, db_size - offset) : 0;
    const size_t dot_chunk = round_up(std::max<size_t>(real, 1), pad);
    if (used[slot]) IRIS_HIP_CHECK(hipStreamWaitEvent(g, freed[slot].e, 0));
    ShareDB* eng[6];
    const QuerySet* qs[6];
    const DbSlice* dbs[6];
    const uint16_t mult[6] = {1, 2, 1, 2, 1, 2};
    for (int i = 0; i < 3; i++) {
      cd[i]->set_slot(slot);
      md[i]->set_slot(slot);
      eng[2 * i] = cd[i];
      qs[2 * i] = p[i].code_query;
      dbs[2 * i] = p[i].code_db;
      eng[2 * i + 1] = md[i];
      qs[2 * i + 1] = p[i].mask_query;
      dbs[2 * i + 1] = p[i].mask_db;
    }
    ShareDB::dot_reduce_joint(eng, qs, dbs, mult, 6, dot_chunk, offset, g);
    ShareDB::reshare_trio(cd, dot_chunk, gs);
    ShareDB::reshare_trio(md, dot_chunk, gs);
    const size_t len = dot_chunk * rows;
    const uint16_t *ca[3], *cb[3], *ma[3], *mb[3];
    for (int i = 0; i < 3; i++) {
      cc[i]->set_chunk_size(len / 64);
      ca[i] = cd[i]->results();
      cb[i] = cd[i]->results_peer();
      ma[i] = md[i]->results();
      mb[i] = md[i]->results_peer();
    }
    Circuits::compare_trio(cc, ca, cb, ma, mb, len, threshold_a, gs, /*fork=*/false);
    Circuits::open_exchange_trio(cc, gs);
    if (real > 0)
      for (int i = 0; i < 3; i++)
        p[i].comparator->open_results(cc[i]->result_a(), cc[i]->result_b(), cc[i]->result_c(),
                                      bitmaps[i], dot_chunk, real, offset, db_size, batch_size, ps,
                                      rows);
    IRIS_HIP_CHECK(hipEventRecord(freed[slot].e, ps));
    used[slot] = true;
    offset += kDbChunk;
    c++;
  } while (offset < db_size);
  Event done;
  IRIS_HIP_CHECK(hipEventRecord(done.e, ps));
  for (int i = 0; i < 3; i++) {
    IRIS_HIP_CHECK(hipStreamWaitEvent(s[i], done.e, 0));
    cd[i]->set_slot(0);
    md[i]->set_slot(0);
  }
}

void scan_db_trio(const PartyScan p[3], size_t db_size, size_t batch_size, uint32_t threshold_a,
                  uint64_t* const bitmaps[3], hipStream_t const s[3]) {
  bool capture = false;
  for (int i = 0; i < 3; i++) capture |= p[i].comparator->anon_capture_enabled();
  if (!capture) {
    scan_db_trio_ordered(p, db_size, batch_size, threshold_a, bitmaps, s);
    return;
  }
  run_trio([&](int i) { scan_db(p[i], db_size, batch_size, threshold_a, bitmaps[i], s[i]); });
}

}  // namespace irismpc
