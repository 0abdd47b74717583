// extern "C" boundary of libirismpc (declared in include/irismpc.h).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/irismpc.h"
#include "actor.hpp"
#include "batching.hpp"
#include "comparator.hpp"
#include "engine.hpp"
#include "host_crypto.hpp"
#include "serving.hpp"
#include "synth.hpp"

using namespace irismpc;

struct irismpc_exchange {
  Exchange impl;
  irismpc_exchange(int d, double t) : impl(d, t) {}
};
struct irismpc_share_db {
  std::unique_ptr<ShareDB> impl;
  int device;
};
struct irismpc_db {
  std::unique_ptr<DbSlice> impl;
};
struct irismpc_query {
  std::unique_ptr<QuerySet> impl;
};
struct irismpc_circuits {
  std::unique_ptr<Circuits> impl;
  int device;
};
struct irismpc_comparator {
  std::unique_ptr<DistanceComparator> impl;
};
struct irismpc_batch_runner {
  std::unique_ptr<BatchRunner> impl;
  int device;
};
struct irismpc_serving {
  std::unique_ptr<Serving> impl;
};

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return -1;
}

ChaChaKey key_of(const uint32_t* s) {
  ChaChaKey k;
  if (s)
    std::memcpy(k.k, s, 32);
  else
    std::memset(k.k, 0, 32);
  return k;
}

hipStream_t st(void* s) { return reinterpret_cast<hipStream_t>(s); }

struct DevSet {
  int prev = 0;
  explicit DevSet(int d) {
    IRIS_HIP_CHECK(hipGetDevice(&prev));
    IRIS_HIP_CHECK(hipSetDevice(d));
  }
  ~DevSet() { (void)hipSetDevice(prev); }
};
}  // namespace

extern "C" {

const char* irismpc_last_error(void) { return g_last_error.c_str(); }
int irismpc_version(void) { return 100; }

int irismpc_device_count(int* n) {
  return guarded([&] { IRIS_HIP_CHECK(hipGetDeviceCount(n)); });
}

int irismpc_stream_create(int device, void** stream) {
  return guarded([&] {
    DevSet g(device);
    hipStream_t s;
    IRIS_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
  });
}
int irismpc_stream_destroy(void* stream) {
  return guarded([&] { IRIS_HIP_CHECK(hipStreamDestroy(st(stream))); });
}
int irismpc_stream_synchronize(void* stream) {
  return guarded([&] { IRIS_HIP_CHECK(hipStreamSynchronize(st(stream))); });
}
int irismpc_malloc(int device, size_t bytes, void** ptr) {
  return guarded([&] {
    DevSet g(device);
    IRIS_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
  });
}
int irismpc_free(void* ptr) {
  return guarded([&] { IRIS_HIP_CHECK(hipFree(ptr)); });
}
int irismpc_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    IRIS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st(stream)));
    IRIS_HIP_CHECK(hipStreamSynchronize(st(stream)));
  });
}
int irismpc_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  return guarded([&] {
    IRIS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st(stream)));
    IRIS_HIP_CHECK(hipStreamSynchronize(st(stream)));
  });
}
int irismpc_fill_u32(void* ptr, size_t n, uint32_t value, void* stream) {
  return guarded([&] { fill_u32(static_cast<uint32_t*>(ptr), n, value, st(stream)); });
}
int irismpc_fill_u64(void* ptr, size_t n, uint64_t value, void* stream) {
  return guarded([&] { fill_u64(static_cast<uint64_t*>(ptr), n, value, st(stream)); });
}
int irismpc_derive_seed(const uint32_t seed[8], uint64_t nonce, uint32_t out[8]) {
  return guarded([&] { derive_seed(seed, nonce, out); });
}

// ------------------------------------------------------------- exchange
int irismpc_exchange_create(int device, double timeout_s, irismpc_exchange** out) {
  return guarded([&] { *out = new irismpc_exchange(device, timeout_s > 0 ? timeout_s : 300.0); });
}
int irismpc_exchange_destroy(irismpc_exchange* ex) {
  return guarded([&] { delete ex; });
}

// ------------------------------------------------------------- ShareDB
int irismpc_share_db_init(int device, int peer_id, size_t max_db_length, size_t query_length,
                          size_t code_length, const uint32_t seed_own[8],
                          const uint32_t seed_prev[8], irismpc_exchange* comm, int tag,
                          irismpc_share_db** out) {
  return guarded([&] {
    auto* h = new irismpc_share_db();
    h->device = device;
    try {
      h->impl = std::make_unique<ShareDB>(device, peer_id, max_db_length, query_length,
                                          (int)code_length, key_of(seed_own), key_of(seed_prev),
                                          comm ? &comm->impl : nullptr, tag);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int irismpc_share_db_destroy(irismpc_share_db* e) {
  return guarded([&] { delete e; });
}
int irismpc_share_db_alloc_db(irismpc_share_db* e, size_t max_db_length, irismpc_db** out) {
  return guarded([&] {
    auto* h = new irismpc_db();
    h->impl = e->impl->alloc_db(max_db_length);
    *out = h;
  });
}
int irismpc_db_destroy(irismpc_db* db) {
  return guarded([&] { delete db; });
}
int irismpc_db_load_records(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                            const uint16_t* records, void* stream) {
  return guarded([&] { e->impl->load_records_u16(*db->impl, index, n, records, st(stream)); });
}
int irismpc_db_load_records_limbs(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                                  const uint8_t* limb0, const uint8_t* limb1, void* stream) {
  return guarded(
      [&] { e->impl->load_records_limbs(*db->impl, index, n, limb0, limb1, st(stream)); });
}
int irismpc_db_write_at_index(irismpc_share_db* e, irismpc_db* db, const irismpc_query* q,
                              size_t src_index, size_t dst_index, void* stream) {
  return guarded(
      [&] { e->impl->write_at_index(*db->impl, *q->impl, src_index, dst_index, st(stream)); });
}
int irismpc_db_load_stored_records(irismpc_share_db* codes, irismpc_share_db* masks,
                                   irismpc_db* left_code, irismpc_db* left_mask,
                                   irismpc_db* right_code, irismpc_db* right_mask,
                                   const uint8_t* records, size_t n_records,
                                   uint32_t max_serial_id, size_t* n_loaded, size_t* row_end,
                                   void* stream) {
  return guarded([&] {
    IRIS_REQUIRE(!n_records || records, "records is NULL");
    DevSet g(codes->device);
    size_t end = 0;
    const size_t n = load_stored_records(*codes->impl, *masks->impl, *left_code->impl,
                                         *left_mask->impl, *right_code->impl, *right_mask->impl,
                                         records, n_records, max_serial_id, &end, st(stream));
    if (n_loaded) *n_loaded = n;
    if (row_end) *row_end = end;
  });
}
int irismpc_db_prefetch_subset(irismpc_share_db* e, const irismpc_db* db, const uint32_t* ids,
                               size_t n, irismpc_db* subset, void* stream) {
  return guarded([&] { e->impl->prefetch_subset(*db->impl, ids, n, *subset->impl, st(stream)); });
}
int irismpc_db_read_records(irismpc_share_db* e, const irismpc_db* db, size_t index, size_t n,
                            uint16_t* out, void* stream) {
  return guarded([&] {
    DevSet g(e->device);
    IRIS_REQUIRE(index + n <= db->impl->capacity, "read beyond DB capacity");
    const int K = e->impl->code_length();
    DevBuf<uint16_t> tmp(n * (size_t)K, false);
    launch_untile_rows(db->impl->limbs.ptr, index, n, K, tmp.ptr, st(stream));
    IRIS_HIP_CHECK(hipMemcpyAsync(out, tmp.ptr, tmp.n * 2, hipMemcpyDeviceToHost, st(stream)));
    IRIS_HIP_CHECK(hipStreamSynchronize(st(stream)));
  });
}
int irismpc_db_fill_random(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                           const uint32_t key[8], uint64_t nonce, void* stream) {
  return guarded([&] {
    DevSet g(e->device);
    IRIS_REQUIRE(index + round_up(n, 16) <= db->impl->capacity, "fill beyond DB capacity");
    launch_fill_random(db->impl->limbs.ptr, index, n, e->impl->code_length(), key_of(key), nonce,
                       db->impl->sum0.ptr, db->impl->sum1.ptr, st(stream));
  });
}
int irismpc_db_load_records_device(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                                   const uint16_t* dev_rows, void* stream) {
  return guarded([&] {
    DevSet g(e->device);
    e->impl->load_records_device(*db->impl, index, n, dev_rows, st(stream));
  });
}
int irismpc_db_replicate_records(irismpc_share_db* e, irismpc_db* db, size_t src, size_t n,
                                 size_t dst, void* stream) {
  return guarded([&] {
    DevSet g(e->device);
    e->impl->replicate_records(*db->impl, src, n, dst, st(stream));
  });
}
struct irismpc_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~irismpc_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};
int irismpc_graph_capture_begin(void* stream) {
  return guarded([&] {
    IRIS_REQUIRE(stream != nullptr, "capture needs a non-default stream");
    IRIS_HIP_CHECK(hipStreamBeginCapture(st(stream), hipStreamCaptureModeThreadLocal));
  });
}
int irismpc_graph_capture_end(void* stream, irismpc_graph** out) {
  return guarded([&] {
    auto g = std::make_unique<irismpc_graph>();
    IRIS_HIP_CHECK(hipStreamEndCapture(st(stream), &g->graph));
    IRIS_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}
int irismpc_graph_launch(irismpc_graph* g, void* stream) {
  return guarded([&] { IRIS_HIP_CHECK(hipGraphLaunch(g->exec, st(stream))); });
}
int irismpc_graph_destroy(irismpc_graph* g) {
  return guarded([&] { delete g; });
}
int irismpc_synth_iris(uint64_t dataset, uint64_t index, uint64_t code[200], uint64_t mask[200]) {
  return guarded([&] { synth_iris_host(dataset, index, code, mask); });
}
int irismpc_synth_flip(uint64_t code[200], uint64_t dataset, uint64_t index, uint32_t n_flips) {
  return guarded([&] { synth_flip_host(code, dataset, index, n_flips); });
}
int irismpc_synth_plain_device(uint64_t dataset, uint64_t index0, size_t n, uint64_t* codes,
                               uint64_t* masks, void* stream) {
  return guarded([&] { launch_synth_plain(dataset, index0, n, codes, masks, st(stream)); });
}
int irismpc_synth_shares_device(const uint64_t* codes, const uint64_t* masks, size_t n,
                                uint64_t rand_dataset, uint64_t rand_index0,
                                uint16_t* const code_out[3], uint16_t* const mask_out[3],
                                void* stream) {
  return guarded([&] {
    launch_synth_shares(codes, masks, n, rand_dataset, rand_index0, code_out, mask_out, st(stream));
  });
}
int irismpc_query_upload(irismpc_share_db* e, const uint16_t* rows, size_t n_rows, void* stream,
                         irismpc_query** out) {
  return guarded([&] {
    auto* h = new irismpc_query();
    h->impl = e->impl->upload_query(rows, n_rows, st(stream));
    *out = h;
  });
}
int irismpc_query_upload_shares(irismpc_share_db* e, const uint16_t* shares, size_t batch,
                                void* stream, irismpc_query** out) {
  return guarded([&] {
    auto* h = new irismpc_query();
    h->impl = e->impl->upload_query_shares(shares, batch, st(stream));
    *out = h;
  });
}
int irismpc_query_upload_shares_rows(irismpc_share_db* e, const uint16_t* shares, size_t batch,
                                     size_t rows, void* stream, irismpc_query** out) {
  return guarded([&] {
    auto* h = new irismpc_query();
    h->impl = e->impl->upload_query_shares_rows(shares, batch, rows, st(stream));
    *out = h;
  });
}
int irismpc_query_refill_shares(irismpc_share_db* e, irismpc_query* q, const uint16_t* shares,
                                size_t batch, size_t rows, void* stream) {
  return guarded([&] { e->impl->refill_query_shares(*q->impl, shares, batch, rows, st(stream)); });
}
int irismpc_query_refill_shares_async(irismpc_share_db* e, irismpc_query* q, const uint16_t* shares,
                                      size_t batch, size_t rows, void* stream) {
  return guarded([&] { e->impl->refill_query_shares_async(*q->impl, shares, batch, rows, st(stream)); });
}
int irismpc_query_upload_shares_mirrored(irismpc_share_db* e, const uint16_t* shares,
                                         const uint16_t* insert_shares, size_t batch,
                                         void* stream, irismpc_query** out) {
  return guarded([&] {
    auto h = std::make_unique<irismpc_query>();
    h->impl = e->impl->upload_query_shares_mirrored(shares, insert_shares, batch, 0, st(stream));
    *out = h.release();
  });
}
int irismpc_query_destroy(irismpc_query* q) {
  return guarded([&] { delete q; });
}
int irismpc_share_db_dot_reduce(irismpc_share_db* e, const irismpc_query* q, const irismpc_db* db,
                                size_t chunk_size, size_t offset, uint16_t multiplier,
                                void* stream) {
  return guarded([&] {
    e->impl->dot_reduce(*q->impl, *db->impl, chunk_size, offset, multiplier, st(stream));
  });
}
int irismpc_share_db_dot_reduce_chunks(irismpc_share_db* e, const irismpc_query* q,
                                       const irismpc_db* db, size_t offset, size_t total,
                                       size_t chunk_size, uint16_t multiplier,
                                       size_t out_offset, void* stream) {
  return guarded([&] {
    e->impl->dot_reduce_chunks(*q->impl, *db->impl, offset, total, chunk_size, multiplier,
                               st(stream), out_offset);
  });
}
int irismpc_share_db_dot_reduce_range(irismpc_share_db* e, const irismpc_query* q,
                                      const irismpc_db* db, size_t offset, size_t total,
                                      size_t chunk_size, uint16_t multiplier, uint16_t* out,
                                      void* stream) {
  return guarded([&] {
    e->impl->dot_reduce_range(*q->impl, *db->impl, offset, total, chunk_size, multiplier, out,
                              st(stream));
  });
}
int irismpc_share_db_dot_reduce_packed(irismpc_share_db* e, const irismpc_query* q,
                                       const irismpc_db* db, size_t offset, size_t total,
                                       size_t chunk_size, uint16_t multiplier,
                                       const size_t* rows, uint16_t* const* outs, int n_batches,
                                       const uint64_t* ctr_first, void* stream) {
  return guarded([&] {
    e->impl->dot_reduce_packed(*q->impl, *db->impl, offset, total, chunk_size, multiplier, rows,
                               outs, n_batches, ctr_first, st(stream));
  });
}
int irismpc_share_db_reshare_results(irismpc_share_db* e, size_t chunk_size, void* stream) {
  return guarded([&] { e->impl->reshare_results(chunk_size, st(stream)); });
}
int irismpc_share_db_reshare_results_trio(irismpc_share_db* const e[3], size_t chunk_size,
                                          void* const streams[3]) {
  return guarded([&] {
    ShareDB* d[3] = {e[0]->impl.get(), e[1]->impl.get(), e[2]->impl.get()};
    hipStream_t s[3] = {st(streams[0]), st(streams[1]), st(streams[2])};
    ShareDB::reshare_trio(d, chunk_size, s);
  });
}
int irismpc_share_db_set_peer_copy(irismpc_share_db* e, int copy) {
  return guarded([&] { e->impl->peer_copy = copy != 0; });
}
int irismpc_share_db_result_views(irismpc_share_db* e, const uint16_t** a, const uint16_t** b) {
  return guarded([&] {
    *a = e->impl->results();
    *b = e->impl->results_peer();
  });
}
int irismpc_share_db_fetch_results(irismpc_share_db* e, uint16_t* out, size_t n, void* stream) {
  return guarded([&] {
    IRIS_REQUIRE(n <= e->impl->results_capacity(), "fetch beyond the result buffer");
    IRIS_HIP_CHECK(hipMemcpyAsync(out, e->impl->results(), n * 2, hipMemcpyDeviceToHost, st(stream)));
    IRIS_HIP_CHECK(hipStreamSynchronize(st(stream)));
  });
}
int irismpc_share_db_get_counter(const irismpc_share_db* e, uint64_t* counter) {
  return guarded([&] { *counter = e->impl->counter(); });
}
int irismpc_share_db_set_counter(irismpc_share_db* e, uint64_t counter) {
  return guarded([&] { e->impl->set_counter(counter); });
}
int irismpc_share_db_set_variant(irismpc_share_db* e, int variant) {
  return guarded([&] {
    if (variant != ShareDB::kAutoVariant) (void)gemm_tile(variant);
    e->impl->variant = variant;
  });
}
int irismpc_share_db_set_timing(irismpc_share_db* e, int on) {
  return guarded([&] { e->impl->timer.enable(on != 0); });
}
int irismpc_share_db_launch_times(irismpc_share_db* e, size_t* launches, double* total_ms) {
  return guarded([&] {
    DevSet d(e->device);
    const size_t n = e->impl->timer.collect(total_ms);
    if (launches) *launches = n;
  });
}
int irismpc_gemm_variants(int* n) {
  return guarded([&] { *n = kGemmVariants; });
}
int irismpc_gemm_rot_variants(int* first, int* n) {
  return guarded([&] {
    *first = kRotVariantBase;
    *n = kRotVariants;
  });
}

// ------------------------------------------------------------- Circuits
int irismpc_circuits_new(int device, int peer_id, size_t input_size, size_t alloc_size,
                         const uint32_t seed_own[8], const uint32_t seed_prev[8],
                         irismpc_exchange* comm, int tag, irismpc_circuits** out) {
  return guarded([&] {
    IRIS_REQUIRE(comm != nullptr, "Circuits needs an exchange handle");
    auto* h = new irismpc_circuits();
    h->device = device;
    try {
      h->impl = std::make_unique<Circuits>(device, peer_id, input_size, alloc_size,
                                           key_of(seed_own), key_of(seed_prev), &comm->impl, tag);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int irismpc_circuits_destroy(irismpc_circuits* c) {
  return guarded([&] { delete c; });
}
int irismpc_circuits_set_chunk_size(irismpc_circuits* c, size_t chunk_words) {
  return guarded([&] { c->impl->set_chunk_size(chunk_words); });
}
int irismpc_circuits_compare_threshold_masked_many(irismpc_circuits* c, const uint16_t* code_a,
                                                   const uint16_t* code_b,
                                                   const uint16_t* mask_a,
                                                   const uint16_t* mask_b, size_t len,
                                                   uint32_t threshold_a, void* stream) {
  return guarded([&] {
    c->impl->compare_threshold_masked_many(code_a, code_b, mask_a, mask_b, len, threshold_a,
                                           st(stream));
  });
}
int irismpc_circuits_compare_threshold_masked_many_trio(
    irismpc_circuits* const c[3], const uint16_t* const code_a[3], const uint16_t* const code_b[3],
    const uint16_t* const mask_a[3], const uint16_t* const mask_b[3], size_t len,
    uint32_t threshold_a, void* const streams[3]) {
  return guarded([&] {
    Circuits* cc[3] = {c[0]->impl.get(), c[1]->impl.get(), c[2]->impl.get()};
    hipStream_t s[3] = {st(streams[0]), st(streams[1]), st(streams[2])};
    Circuits::compare_trio(cc, code_a, code_b, mask_a, mask_b, len, threshold_a, s);
  });
}
int irismpc_circuits_result_buffer(irismpc_circuits* c, const uint64_t** a, const uint64_t** b) {
  return guarded([&] {
    *a = c->impl->result_a();
    *b = c->impl->result_b();
  });
}
int irismpc_circuits_open(irismpc_circuits* c, void* stream) {
  return guarded([&] { c->impl->open_exchange(st(stream)); });
}
int irismpc_circuits_open_trio(irismpc_circuits* const c[3], void* const streams[3]) {
  return guarded([&] {
    Circuits* cc[3] = {c[0]->impl.get(), c[1]->impl.get(), c[2]->impl.get()};
    hipStream_t s[3] = {st(streams[0]), st(streams[1]), st(streams[2])};
    Circuits::open_exchange_trio(cc, s);
  });
}
int irismpc_circuits_open_views(irismpc_circuits* c, const uint64_t** a, const uint64_t** b,
                                const uint64_t** cc) {
  return guarded([&] {
    *a = c->impl->result_a();
    *b = c->impl->result_b();
    *cc = c->impl->result_c();
  });
}
int irismpc_circuits_get_counter(const irismpc_circuits* c, uint64_t* counter) {
  return guarded([&] { *counter = c->impl->counter(); });
}
int irismpc_circuits_set_counter(irismpc_circuits* c, uint64_t counter) {
  return guarded([&] { c->impl->set_counter(counter); });
}
int irismpc_circuits_set_timing(irismpc_circuits* c, int on) {
  return guarded([&] { c->impl->timer.enable(on != 0); });
}
int irismpc_circuits_launch_times(irismpc_circuits* c, size_t* launches, double* total_ms) {
  return guarded([&] {
    DevSet d(c->device);
    const size_t n = c->impl->timer.collect(total_ms);
    if (launches) *launches = n;
  });
}
int irismpc_circuits_set_lift_capture(irismpc_circuits* c, uint32_t* a, uint32_t* b) {
  return guarded([&] {
    c->impl->lifted_dbg_a = a;
    c->impl->lifted_dbg_b = b;
  });
}

// ------------------------------------------------------------- DistanceComparator
int irismpc_comparator_init(int device, size_t query_length, size_t max_db_size,
                            irismpc_comparator** out) {
  return guarded([&] {
    auto* h = new irismpc_comparator();
    try {
      h->impl = std::make_unique<DistanceComparator>(device, query_length, max_db_size);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}
int irismpc_comparator_destroy(irismpc_comparator* d) {
  return guarded([&] { delete d; });
}
int irismpc_comparator_open_results(irismpc_comparator* d, const uint64_t* a, const uint64_t* b,
                                    const uint64_t* c, uint64_t* bitmap, size_t chunk_len,
                                    size_t real_chunk_len, size_t offset, size_t total_db_len,
                                    size_t batch_size, void* stream) {
  return guarded([&] {
    d->impl->open_results(a, b, c, bitmap, chunk_len, real_chunk_len, offset, total_db_len,
                          batch_size, st(stream));
  });
}
int irismpc_comparator_join_db_matches(irismpc_comparator* d, const uint64_t* left,
                                       const uint64_t* right, uint32_t* final_results,
                                       size_t db_size, const uint64_t* or_policy,
                                       size_t max_db_size, void* stream) {
  return guarded([&] {
    d->impl->join_db_matches(left, right, final_results, db_size, or_policy, max_db_size,
                             st(stream));
  });
}
int irismpc_comparator_get_partial_results(irismpc_comparator* d, const uint64_t* bitmap,
                                           size_t db_size, size_t batch_size, size_t row_bits,
                                           const uint32_t* extra_ids, size_t n_extra,
                                           uint32_t* host_ids, size_t host_capacity,
                                           size_t* n_ids, void* stream) {
  return guarded([&] {
    IRIS_REQUIRE(n_ids, "n_ids is NULL");
    IRIS_REQUIRE(!n_extra || extra_ids, "extra_ids is NULL");
    const size_t n = d->impl->get_partial_results(bitmap, db_size, batch_size, extra_ids, n_extra,
                                                  st(stream), row_bits);
    *n_ids = n;
    const size_t m = std::min(n, host_capacity);
    if (host_ids && m) {
      IRIS_HIP_CHECK(hipMemcpyAsync(host_ids, d->impl->partial_ids(), m * 4,
                                    hipMemcpyDeviceToHost, st(stream)));
      IRIS_HIP_CHECK(hipStreamSynchronize(st(stream)));
    }
  });
}
int irismpc_comparator_join_batch_matches(irismpc_comparator* d, const uint64_t* left,
                                          const uint64_t* right, uint32_t* final_results,
                                          size_t batch_size, size_t row_bits, void* stream) {
  return guarded([&] {
    d->impl->join_batch_matches(left, right, final_results, batch_size, row_bits, st(stream));
  });
}
int irismpc_comparator_set_anon_capture(irismpc_comparator* d, uint16_t* code_a, uint16_t* code_b,
                                        uint16_t* mask_a, uint16_t* mask_b, uint64_t* idx,
                                        uint32_t* counter, size_t capacity, uint64_t batch_id,
                                        size_t max_query_length, size_t max_db_length,
                                        const uint64_t* targets, size_t n_targets) {
  return guarded([&] {
    AnonCapture c;
    c.code_a = code_a; c.code_b = code_b; c.mask_a = mask_a; c.mask_b = mask_b;
    c.idx = idx; c.counter = counter; c.capacity = capacity; c.batch_id = batch_id;
    c.max_query_length = max_query_length; c.max_db_length = max_db_length;
    d->impl->set_anon_capture(c, targets, n_targets);
  });
}
int irismpc_partial_results_noalloc(const uint64_t* bitmap, size_t db_size, size_t batch_size,
                                    size_t row_bits, const uint32_t* extra_ids, size_t n_extra,
                                    uint32_t* ids, uint32_t* count, void* stream) {
  return guarded([&] {
    partial_results_noalloc(bitmap, db_size, batch_size, row_bits, extra_ids, n_extra, ids, count,
                            st(stream));
  });
}
int irismpc_comparator_partial_ids(irismpc_comparator* d, const uint32_t** ids) {
  return guarded([&] { *ids = d->impl->partial_ids(); });
}
int irismpc_comparator_open_results_with_index_mapping(
    irismpc_comparator* d, const uint64_t* a, const uint64_t* b, const uint64_t* c,
    uint64_t* bitmap, size_t chunk_len, size_t real_chunk_len, size_t total_db_len,
    size_t batch_size, const uint32_t* index_map, void* stream) {
  return guarded([&] {
    d->impl->open_results_mapped(a, b, c, bitmap, chunk_len, real_chunk_len, total_db_len,
                                 batch_size, index_map, st(stream));
  });
}
int irismpc_comparator_fetch_partial_rotations(irismpc_comparator* d, uint32_t* query,
                                               uint32_t* db, int8_t* rotation, size_t capacity,
                                               size_t* n_matches, void* stream) {
  return guarded([&] {
    IRIS_REQUIRE(n_matches, "n_matches is NULL");
    IRIS_REQUIRE(!capacity || (query && db && rotation), "capture arrays are NULL");
    *n_matches = d->impl->fetch_partial_rotations(query, db, rotation, capacity, st(stream));
  });
}
int irismpc_comparator_reset(irismpc_comparator* d, void* stream) {
  return guarded([&] { d->impl->reset(st(stream)); });
}
int irismpc_comparator_buffers(irismpc_comparator* d, uint32_t** match_counters,
                               uint32_t** all_matches, uint32_t** counters_left,
                               uint32_t** counters_right, uint32_t** partial_left,
                               uint32_t** partial_right) {
  return guarded([&] {
    *match_counters = d->impl->match_counters();
    *all_matches = d->impl->all_matches();
    *counters_left = d->impl->match_counters_left();
    *counters_right = d->impl->match_counters_right();
    *partial_left = d->impl->partial_results_left();
    *partial_right = d->impl->partial_results_right();
  });
}

// ------------------------------------------------------------- DB scan (actor chunk loop)
static PartyScan party_scan(irismpc_share_db* codes, irismpc_share_db* masks,
                            irismpc_circuits* circuits, irismpc_comparator* cmp,
                            const irismpc_db* code_db, const irismpc_db* mask_db,
                            const irismpc_query* q_code, const irismpc_query* q_mask) {
  PartyScan p{};
  p.codes = codes->impl.get();
  p.masks = masks->impl.get();
  p.circuits = circuits->impl.get();
  p.comparator = cmp->impl.get();
  p.code_db = code_db->impl.get();
  p.mask_db = mask_db->impl.get();
  p.code_query = q_code->impl.get();
  p.mask_query = q_mask->impl.get();
  return p;
}

int irismpc_scan_db(irismpc_share_db* codes, irismpc_share_db* masks, irismpc_circuits* circuits,
                    irismpc_comparator* cmp, const irismpc_db* code_db, const irismpc_db* mask_db,
                    const irismpc_query* q_code, const irismpc_query* q_mask, size_t db_size,
                    size_t batch_size, uint32_t threshold_a, uint64_t* bitmap, void* stream) {
  return guarded([&] {
    scan_db(party_scan(codes, masks, circuits, cmp, code_db, mask_db, q_code, q_mask), db_size,
            batch_size, threshold_a, bitmap, st(stream));
  });
}

int irismpc_scan_db_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                         irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                         const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                         const irismpc_query* const q_code[3], const irismpc_query* const q_mask[3],
                         size_t db_size, size_t batch_size, uint32_t threshold_a,
                         uint64_t* const bitmaps[3], void* const streams[3]) {
  return guarded([&] {
    PartyScan p[3];
    hipStream_t s[3];
    for (int i = 0; i < 3; i++) {
      p[i] = party_scan(codes[i], masks[i], circuits[i], cmp[i], code_db[i], mask_db[i], q_code[i],
                        q_mask[i]);
      s[i] = st(streams[i]);
    }
    scan_db_trio(p, db_size, batch_size, threshold_a, bitmaps, s);
  });
}
int irismpc_scan_chunk_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                            irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                            const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                            const irismpc_query* const q_code[3], const irismpc_query* const q_mask[3],
                            size_t offset, size_t db_size, size_t batch_size, uint32_t threshold_a,
                            uint64_t* const bitmaps[3], const uint32_t* const slot_of[3],
                            size_t row_bits, void* const streams[3]) {
  return guarded([&] {
    PartyScan p[3];
    hipStream_t s[3];
    for (int i = 0; i < 3; i++) {
      p[i] = party_scan(codes[i], masks[i], circuits[i], cmp[i], code_db[i], mask_db[i], q_code[i],
                        q_mask[i]);
      s[i] = st(streams[i]);
    }
    scan_chunk_trio(p, offset, db_size, batch_size, threshold_a, bitmaps, slot_of, row_bits, s);
  });
}

int irismpc_scan_chunks_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                             irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                             const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                             const irismpc_query* const q_code[3], const irismpc_query* const q_mask[3],
                             size_t offset, size_t n_chunks, size_t db_size, size_t batch_size,
                             uint32_t threshold_a, uint64_t* const bitmaps[3],
                             const uint32_t* const slot_of[3], size_t row_bits, void* const streams[3]) {
  return guarded([&] {
    PartyScan p[3];
    hipStream_t s[3];
    for (int i = 0; i < 3; i++) {
      p[i] = party_scan(codes[i], masks[i], circuits[i], cmp[i], code_db[i], mask_db[i], q_code[i],
                        q_mask[i]);
      s[i] = st(streams[i]);
    }
    scan_chunks_trio(p, offset, n_chunks, db_size, batch_size, threshold_a, bitmaps, slot_of, row_bits, s);
  });
}

int irismpc_scan_subset_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                             irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                             const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                             irismpc_db* const code_sub[3], irismpc_db* const mask_sub[3],
                             const irismpc_query* const q_code[3],
                             const irismpc_query* const q_mask[3], const uint32_t* const ids[3],
                             size_t n_ids, size_t total_db_len, size_t batch_size,
                             uint32_t threshold_a, uint32_t anon_threshold,
                             uint64_t* const bitmaps[3], void* const streams[3]) {
  return guarded([&] {
    PartyScan p[3];
    hipStream_t s[3];
    DbSlice* cs[3];
    DbSlice* ms[3];
    for (int i = 0; i < 3; i++) {
      p[i] = party_scan(codes[i], masks[i], circuits[i], cmp[i], code_db[i], mask_db[i], q_code[i],
                        q_mask[i]);
      s[i] = st(streams[i]);
      cs[i] = code_sub[i]->impl.get();
      ms[i] = mask_sub[i]->impl.get();
    }
    scan_subset_trio(p, cs, ms, ids, n_ids, total_db_len, batch_size, threshold_a, bitmaps, s,
                     anon_threshold);
  });
}

int irismpc_scan_batch_self_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                                 irismpc_circuits* const circuits[3],
                                 irismpc_comparator* const cmp[3],
                                 const irismpc_query* const q_code[3],
                                 const irismpc_query* const q_mask[3], size_t batch_size,
                                 uint32_t threshold_a, uint64_t* const bitmaps[3], size_t row_bits,
                                 void* const streams[3]) {
  return guarded([&] {
    PartyScan p[3];
    hipStream_t s[3];
    for (int i = 0; i < 3; i++) {
      p[i] = PartyScan{};
      p[i].codes = codes[i]->impl.get();
      p[i].masks = masks[i]->impl.get();
      p[i].circuits = circuits[i]->impl.get();
      p[i].comparator = cmp[i]->impl.get();
      p[i].code_query = q_code[i]->impl.get();
      p[i].mask_query = q_mask[i]->impl.get();
      s[i] = st(streams[i]);
    }
    scan_batch_self_trio(p, batch_size, threshold_a, bitmaps, row_bits, s);
  });
}

// ------------------------------------------------------------------ batching form
int irismpc_batch_plan_packs(const uint32_t* sizes, size_t n_batches, size_t max_queries,
                             size_t max_batches, uint32_t* pack_of_batch, size_t* n_packs) {
  return guarded([&] {
    const auto packs = plan_packs(sizes, n_batches, max_queries, max_batches);
    for (size_t p = 0; p < packs.size(); p++)
      for (size_t i : packs[p]) pack_of_batch[i] = (uint32_t)p;
    *n_packs = packs.size();
  });
}

int irismpc_batch_counter_advance(size_t total, size_t chunk, size_t rows, uint64_t* blocks) {
  return guarded([&] {
    IRIS_REQUIRE(chunk > 0, "chunk of 0 rows");
    *blocks = ShareDB::range_counter_advance(total, chunk, rows);
  });
}

int irismpc_batch_scan_counter_advance(size_t n_rows, size_t chunk, size_t rows, uint64_t* blocks) {
  return guarded([&] {
    IRIS_REQUIRE(chunk > 0, "chunk of 0 rows");
    *blocks = scan_counter_advance(n_rows, chunk, rows);
  });
}

int irismpc_batch_runner_create(irismpc_share_db* e, const irismpc_db* db, size_t n_rows,
                                uint16_t multiplier, const uint32_t* sizes, size_t n_batches,
                                const uint16_t* shares, uint16_t* const* outs, size_t chunk,
                                size_t max_queries, size_t max_batches, void* stream,
                                irismpc_batch_runner** out) {
  return guarded([&] {
    DevSet g(e->device);
    auto r = std::make_unique<irismpc_batch_runner>();
    r->device = e->device;
    r->impl = std::make_unique<BatchRunner>(e->impl.get(), db->impl.get(), n_rows, multiplier, sizes, n_batches,
                                            shares, outs, chunk, max_queries, max_batches, st(stream));
    *out = r.release();
  });
}

int irismpc_batch_runner_destroy(irismpc_batch_runner* r) {
  return guarded([&] { delete r; });
}

int irismpc_batch_runner_replan(irismpc_batch_runner* r, size_t max_queries, void* stream) {
  return guarded([&] {
    DevSet g(r->device);
    r->impl->replan(max_queries, st(stream));
  });
}

int irismpc_batch_runner_plan(const irismpc_batch_runner* r, uint32_t* pack_of_batch, size_t* n_packs) {
  return guarded([&] {
    const auto& packs = r->impl->plan();
    for (size_t p = 0; p < packs.size(); p++)
      for (size_t i : packs[p]) pack_of_batch[i] = (uint32_t)p;
    *n_packs = packs.size();
  });
}

int irismpc_batch_runner_issue_packed(irismpc_batch_runner* r, uint64_t ctr0, size_t n_rows, void* stream) {
  return guarded([&] {
    DevSet g(r->device);
    r->impl->issue_packed(ctr0, n_rows, st(stream));
  });
}

int irismpc_batch_runner_prepare_persistent(irismpc_batch_runner* r, uint64_t ctr0, void* stream) {
  return guarded([&] {
    DevSet g(r->device);
    r->impl->prepare_persistent(ctr0, st(stream));
  });
}

int irismpc_batch_runner_issue_persistent(irismpc_batch_runner* r, void* stream) {
  return guarded([&] {
    DevSet g(r->device);
    r->impl->issue_persistent(st(stream));
  });
}

int irismpc_batch_runner_issue_plain(irismpc_batch_runner* r, size_t batch, void* stream) {
  return guarded([&] {
    DevSet g(r->device);
    r->impl->issue_plain(batch, st(stream));
  });
}

// ------------------------------------------------------------------ serving form
namespace {
void copy_matches(const std::vector<uint32_t>& m, uint32_t* matches, size_t max_matches, size_t* n_matches) {
  if (matches) std::copy_n(m.begin(), std::min(max_matches, m.size()), matches);
  if (n_matches) *n_matches = m.size();
}
}  // namespace

int irismpc_serving_create(int device, const uint32_t* party_seeds, irismpc_db* const code_dbs[3],
                           irismpc_db* const mask_dbs[3], size_t n_rows, size_t capacity,
                           size_t max_active, uint32_t threshold_a, irismpc_serving** out) {
  return guarded([&] {
    DevSet g(device);
    ChaChaKey seeds[3];
    DbSlice* cd[3];
    DbSlice* md[3];
    for (int p = 0; p < 3; p++) {
      seeds[p] = key_of(party_seeds + 8 * p);
      IRIS_REQUIRE(code_dbs[p] && mask_dbs[p], "three code and three mask DB slices");
      cd[p] = code_dbs[p]->impl.get();
      md[p] = mask_dbs[p]->impl.get();
    }
    auto s = std::make_unique<irismpc_serving>();
    s->impl = std::make_unique<Serving>(device, seeds, cd, md, n_rows, capacity, max_active, threshold_a);
    *out = s.release();
  });
}

int irismpc_serving_destroy(irismpc_serving* s) {
  return guarded([&] { delete s; });
}

int irismpc_serving_start(irismpc_serving* s, int policy, size_t max_batch, size_t reserve_bytes) {
  return guarded([&] { s->impl->start(policy, max_batch, reserve_bytes); });
}

int irismpc_serving_finish(irismpc_serving* s, double timeout_s) {
  return guarded([&] { s->impl->finish(timeout_s); });
}

int irismpc_serving_submit(irismpc_serving* s, int kind, const uint16_t* code_shares,
                           const uint16_t* mask_shares, int64_t target, double arrival_s, uint64_t* id) {
  return guarded([&] { *id = s->impl->submit(kind, code_shares, mask_shares, target, arrival_s); });
}

int irismpc_serving_result(const irismpc_serving* s, uint64_t id, int* state, int64_t* value,
                           uint32_t* matches, size_t max_matches, size_t* n_matches, double* times,
                           size_t* db_rows, char* error, size_t error_len) {
  return guarded([&] {
    Serving::Result r;
    IRIS_REQUIRE(s->impl->result(id, &r), "unknown request id");
    if (state) *state = r.state;
    if (value) *value = r.value;
    copy_matches(r.matches, matches, max_matches, n_matches);
    if (times) {
      times[0] = r.arrival;
      times[1] = r.done;
    }
    if (db_rows) *db_rows = r.db_rows;
    if (error && error_len) {
      const size_t n = std::min(error_len - 1, r.error.size());
      std::memcpy(error, r.error.data(), n);
      error[n] = 0;
    }
  });
}

int irismpc_serving_result_version(const irismpc_serving* s, uint64_t id, uint64_t* db_version) {
  return guarded([&] {
    Serving::Result r;
    IRIS_REQUIRE(s->impl->result(id, &r), "unknown request id");
    IRIS_REQUIRE(r.state != kReqPending, "request still pending");
    *db_version = r.db_version;
  });
}

int irismpc_serving_result_batch_matches(const irismpc_serving* s, uint64_t id, uint64_t* ids, size_t max_ids,
                                         size_t* n_ids) {
  return guarded([&] {
    Serving::Result r;
    IRIS_REQUIRE(s->impl->result(id, &r), "unknown request id");
    IRIS_REQUIRE(r.state != kReqPending, "request still pending");
    if (ids) std::copy_n(r.batch_matches.begin(), std::min(max_ids, r.batch_matches.size()), ids);
    if (n_ids) *n_ids = r.batch_matches.size();
  });
}

int irismpc_serving_release(irismpc_serving* s, uint64_t id) {
  return guarded([&] { IRIS_REQUIRE(s->impl->release(id), "request still pending"); });
}

int irismpc_serving_now(const irismpc_serving* s, double* seconds) {
  return guarded([&] { *seconds = s->impl->now(); });
}

int irismpc_serving_stats(const irismpc_serving* s, uint64_t* out) {
  return guarded([&] {
    const ServingStats st = s->impl->stats();
    out[0] = st.steps;
    out[1] = st.max_concurrent;
    out[2] = st.admission_waits;
    out[3] = st.batches;
    out[4] = st.refills;
  });
}

int irismpc_serving_loop_times(const irismpc_serving* s, double* out) {
  return guarded([&] {
    const ServingStats st = s->impl->stats();
    out[0] = st.t_refill;
    out[1] = st.t_wait;
    out[2] = st.t_enqueue;
    out[3] = st.t_admit;
    out[4] = st.t_decode;
  });
}

int irismpc_serving_n_rows(const irismpc_serving* s, size_t* n) {
  return guarded([&] { *n = s->impl->n_rows(); });
}

int irismpc_serving_run_alone(irismpc_serving* s, const uint16_t* code_shares, const uint16_t* mask_shares,
                              uint32_t* matches, size_t max_matches, size_t* n_matches) {
  return guarded([&] {
    IRIS_REQUIRE(code_shares && mask_shares, "request without shares");
    copy_matches(s->impl->run_alone(code_shares, mask_shares), matches, max_matches, n_matches);
  });
}

int irismpc_serving_time_step(irismpc_serving* s, size_t batch, size_t n_chunks, int reps, double* seconds) {
  return guarded([&] { *seconds = s->impl->time_step(batch, n_chunks, reps); });
}

int irismpc_serving_set_admission(irismpc_serving* s, size_t active_cap, size_t queue_threshold) {
  return guarded([&] { s->impl->set_admission(active_cap, queue_threshold); });
}

int irismpc_serving_set_step_table(irismpc_serving* s, const uint32_t* buckets, const uint32_t* chunks,
                                   size_t n) {
  return guarded([&] { s->impl->set_step_table(buckets, chunks, n); });
}

int irismpc_serving_tune_steps(irismpc_serving* s, const uint32_t* buckets, size_t n_buckets,
                               const uint32_t* ks, size_t n_ks, double tol, double max_step_s,
                               uint32_t* chunks_out, double* seconds_per_chunk_out) {
  return guarded([&] {
    Serving& sv = *s->impl;
    const size_t n_chunks = Serving::n_chunks_of(sv.n_rows());
    std::vector<uint32_t> bk, ch;
    for (size_t b = 0; b < n_buckets; b++) {
      chunks_out[b] = 0;
      for (size_t j = 0; j < n_ks; j++) seconds_per_chunk_out[b * n_ks + j] = -1;
      if (buckets[b] < 1 || buckets[b] > sv.max_active()) continue;
      std::vector<uint32_t> kk;
      std::vector<double> tt;
      for (size_t j = 0; j < n_ks; j++)
        if (ks[j] >= 1 && (size_t)ks[j] * buckets[b] <= sv.max_active() && ks[j] <= n_chunks &&
            ks[j] <= kMaxStepChunks) {
          const double t = sv.time_step(buckets[b], ks[j], 2) / ks[j];
          seconds_per_chunk_out[b * n_ks + j] = t;
          kk.push_back(ks[j]);
          tt.push_back(t);
        }
      if (kk.empty()) {
        kk.push_back(1);
        tt.push_back(sv.time_step(buckets[b], 1, 2));
      }
      chunks_out[b] = pick_step_chunks(kk.data(), tt.data(), kk.size(), tol, max_step_s);
      bk.push_back(buckets[b]);
      ch.push_back(chunks_out[b]);
    }
    sv.set_step_table(bk.data(), ch.data(), bk.size());
  });
}

int irismpc_serving_chunks_per_step(const irismpc_serving* s, size_t batch, size_t chunk_index,
                                    size_t n_chunks, size_t* k) {
  return guarded([&] { *k = s->impl->chunks_per_step(batch, chunk_index, n_chunks); });
}

int irismpc_serving_pick_step(const uint32_t* ks, const double* seconds_per_chunk, size_t n, double tol,
                              double max_step_s, uint32_t* k) {
  return guarded([&] { *k = pick_step_chunks(ks, seconds_per_chunk, n, tol, max_step_s); });
}

}  // extern "C"
