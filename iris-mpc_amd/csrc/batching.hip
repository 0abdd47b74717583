// Batching form: pack planner and the packed / plain schedules (host code
// over ShareDB's share-GEMM launches).
#include "batching.hpp"

#include <algorithm>
#include <limits>

namespace irismpc {

namespace {
// [offset, padded width) of each chunk of a scan over n_rows rows
std::vector<std::pair<size_t, size_t>> chunk_ranges(size_t n_rows, size_t chunk) {
  std::vector<std::pair<size_t, size_t>> r;
  for (size_t off = 0; off < n_rows; off += chunk) r.emplace_back(off, round_up(std::min(chunk, n_rows - off), 64));
  return r;
}
}  // namespace

std::vector<std::vector<size_t>> plan_packs(const uint32_t* sizes, size_t n, size_t max_queries,
                                            size_t max_batches) {
  IRIS_REQUIRE(max_queries >= 1 && max_batches >= 1, "empty packs");
  std::vector<std::vector<size_t>> packs;
  std::vector<size_t> cur;
  size_t q = 0;
  for (size_t i = 0; i < n; i++) {
    if (!cur.empty() && (q + sizes[i] > max_queries || cur.size() == max_batches)) {
      packs.push_back(cur);
      cur.clear();
      q = 0;
    }
    cur.push_back(i);
    q += sizes[i];
  }
  if (!cur.empty()) packs.push_back(cur);
  return packs;
}

uint64_t scan_counter_advance(size_t n_rows, size_t chunk, size_t rows) {
  uint64_t n = 0;
  for (const auto& r : chunk_ranges(n_rows, chunk)) n += ShareDB::range_counter_advance(r.second, chunk, rows);
  return n;
}

uint32_t pick_step_chunks(const uint32_t* k, const double* t, size_t n, double tol, double max_step_s) {
  IRIS_REQUIRE(n >= 1, "no step sizes timed");
  double best = std::numeric_limits<double>::infinity();
  bool any = false;
  for (size_t i = 0; i < n; i++)
    if (max_step_s <= 0 || k[i] * t[i] <= max_step_s) {
      best = std::min(best, t[i]);
      any = true;
    }
  uint32_t pick = std::numeric_limits<uint32_t>::max();
  for (size_t i = 0; i < n; i++) {
    const bool fits = max_step_s <= 0 || k[i] * t[i] <= max_step_s;
    if (!any || (fits && t[i] <= best * (1 + tol))) pick = std::min(pick, k[i]);
  }
  return pick;
}

BatchRunner::BatchRunner(ShareDB* engine, const DbSlice* db, size_t n_rows, uint16_t multiplier,
                         const uint32_t* sizes, size_t n_batches, const uint16_t* shares,
                         uint16_t* const* outs, size_t chunk, size_t max_queries, size_t max_batches,
                         hipStream_t s)
    : e_(engine),
      db_(db),
      n_rows_(n_rows),
      chunk_(chunk),
      max_queries_(max_queries),
      max_batches_(max_batches),
      mult_(multiplier),
      sizes_(sizes, sizes + n_batches),
      outs_(outs, outs + n_batches) {
  IRIS_REQUIRE(engine && db && db->code_length == engine->code_length(), "DB of another code length");
  IRIS_REQUIRE(n_batches >= 1 && n_rows >= 1 && chunk >= 256 && chunk % 256 == 0, "bad batching geometry");
  IRIS_REQUIRE(n_rows <= db->capacity - kMaxBD, "scan beyond the DB");
  const size_t K = (size_t)engine->code_length();
  size_t q = 0;
  for (size_t i = 0; i < n_batches; i++) {
    IRIS_REQUIRE(sizes[i] >= 1 && outs[i], "empty batch or missing output");
    first_.push_back(q);
    q += sizes[i];
  }
  shares_.assign(shares, shares + q * K);
  for (size_t i = 0; i < n_batches; i++)
    q_plain_.push_back(e_->upload_query_shares_rows(shares_.data() + first_[i] * K, sizes_[i],
                                                    (size_t)sizes_[i] * kRotations, s));
  replan(max_queries, s);
}

void BatchRunner::replan(size_t max_queries, hipStream_t s) {
  plan_ = plan_packs(sizes_.data(), sizes_.size(), max_queries, max_batches_);
  max_queries_ = max_queries;
  q_packed_.clear();  // the old sets are freed before the new ones exist
  n_ranges_ = 0;      // persistent descriptors point at the old sets
  const size_t K = (size_t)e_->code_length();
  for (const auto& pack : plan_) {
    // a pack's batches are consecutive: their shares sit back to back
    size_t nq = 0;
    for (size_t i : pack) nq += sizes_[i];
    q_packed_.push_back(
        e_->upload_query_shares_rows(shares_.data() + first_[pack.front()] * K, nq, nq * kRotations, s));
  }
}

void BatchRunner::issue_packed(uint64_t ctr0, size_t n_rows, hipStream_t s) {
  const size_t n = n_rows ? std::min(n_rows, n_rows_) : n_rows_;
  const bool remote = e_->is_remote();
  std::vector<uint64_t> base(sizes_.size());
  uint64_t c = ctr0;
  for (size_t i = 0; i < sizes_.size(); i++) {  // the plain form's counter at the start of each batch
    base[i] = c;
    if (remote) c += scan_counter_advance(n, chunk_, (size_t)sizes_[i] * kRotations);
  }
  const auto ranges = chunk_ranges(n, chunk_);
  std::vector<size_t> rows;
  std::vector<uint16_t*> outs;
  std::vector<uint64_t> ctr;
  for (size_t p = 0; p < plan_.size(); p++) {
    rows.clear();
    outs.clear();
    for (size_t i : plan_[p]) {
      rows.push_back((size_t)sizes_[i] * kRotations);
      outs.push_back(outs_[i]);
    }
    for (const auto& r : ranges) {
      ctr.clear();
      for (size_t j = 0; j < plan_[p].size(); j++)
        ctr.push_back(base[plan_[p][j]] + (remote ? ShareDB::range_counter_advance(r.first, chunk_, rows[j]) : 0));
      e_->dot_reduce_packed(*q_packed_[p], *db_, r.first, r.second, chunk_, mult_, rows.data(), outs.data(),
                            (int)rows.size(), ctr.data(), s);
    }
  }
}

void BatchRunner::prepare_persistent(uint64_t ctr0, hipStream_t s) {
  const bool remote = e_->is_remote();
  IRIS_REQUIRE(plan_.size() <= (size_t)kMaxSegments, "at most 64 packs per persistent launch");
  std::vector<uint64_t> base(sizes_.size());
  uint64_t c = ctr0;
  for (size_t i = 0; i < sizes_.size(); i++) {
    base[i] = c;
    if (remote) c += scan_counter_advance(n_rows_, chunk_, (size_t)sizes_[i] * kRotations);
  }
  const auto ranges = chunk_ranges(n_rows_, chunk_);
  const size_t P = plan_.size();
  std::vector<GemmArgs> host(ranges.size() * P);
  std::vector<uint32_t> ends(ranges.size() * P);
  std::vector<size_t> rows;
  std::vector<uint16_t*> outs;
  std::vector<uint64_t> ctr;
  const GemmTile t = gemm_tile(kRotVariantBase);
  for (size_t r = 0; r < ranges.size(); r++) {
    uint64_t tiles = 0;
    for (size_t p = 0; p < P; p++) {
      rows.clear();
      outs.clear();
      ctr.clear();
      for (size_t i : plan_[p]) {
        rows.push_back((size_t)sizes_[i] * kRotations);
        outs.push_back(outs_[i]);
        ctr.push_back(base[i] + (remote ? ShareDB::range_counter_advance(ranges[r].first, chunk_, rows.back()) : 0));
      }
      IRIS_REQUIRE(q_packed_[p]->rot.ptr && (size_t)e_->code_length() % 800 == 0,
                   "the persistent form needs rotation-structured query sets");
      GemmArgs& a = host[r * P + p];
      a = e_->packed_args(*q_packed_[p], *db_, ranges[r].first, ranges[r].second, chunk_, mult_, rows.data(),
                          outs.data(), (int)rows.size(), ctr.data(), kRotVariantBase);
      IRIS_REQUIRE(a.n_qtiles == (int)ceil_div((size_t)a.qr, (size_t)t.bq), "descriptor tile geometry");
      tiles += (uint64_t)a.n_qtiles * (uint64_t)a.n_dtiles;
      IRIS_REQUIRE(tiles < (1ull << 31), "too many tiles for one persistent launch");
      ends[r * P + p] = (uint32_t)tiles;
    }
  }
  int cus = 0;
  IRIS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e_->device()));
  grid_ = (int)std::max<size_t>(8, round_up((size_t)std::max(cus, 8), 8));
  descs_.alloc(host.size(), false);
  tile_end_.alloc(ends.size(), false);
  IRIS_HIP_CHECK(hipMemcpyAsync(descs_.ptr, host.data(), host.size() * sizeof(GemmArgs), hipMemcpyHostToDevice, s));
  IRIS_HIP_CHECK(hipMemcpyAsync(tile_end_.ptr, ends.data(), ends.size() * 4, hipMemcpyHostToDevice, s));
  IRIS_HIP_CHECK(hipStreamSynchronize(s));
  n_ranges_ = ranges.size();
}

void BatchRunner::issue_persistent(hipStream_t s) {
  IRIS_REQUIRE(n_ranges_ > 0, "persistent launches need prepare_persistent first");
  const size_t P = plan_.size();
  for (size_t r = 0; r < n_ranges_; r++)
    launch_share_gemm_multi(descs_.ptr + r * P, tile_end_.ptr + r * P, (int)P, grid_, s);
}

void BatchRunner::issue_plain(size_t batch, hipStream_t s) {
  IRIS_REQUIRE(batch < sizes_.size(), "batch index beyond the batches");
  for (const auto& r : chunk_ranges(n_rows_, chunk_))
    e_->dot_reduce_range(*q_plain_[batch], *db_, r.first, r.second, chunk_, mult_, outs_[batch], s);
}

}  // namespace irismpc
