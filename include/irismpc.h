/*
 * irismpc: MI355X (gfx950) implementation of the iris-mpc GPU server's
 * secret-shared matching path, behind a C ABI.
 *
 * Every entry point returns 0 on success and a negative value on failure; the
 * failure text is available from irismpc_last_error() on the same thread.
 * Nothing aborts across the boundary. Device buffers are plain pointers
 * (hipStream_t / device pointers passed as void*), sizes are element counts
 * unless stated otherwise.
 *
 * Party model: the three semi-honest parties run on one GPU. Each party owns
 * its ShareDB / Circuits engines and its own stream. The NcclComm of the
 * reference is replaced by an irismpc_exchange shared by the three parties:
 * calls the reference makes as matched NCCL send/recv pairs (reshare, the
 * circuit rounds, open) are collectives - every party calls with its own
 * handle and stream (from its own thread, exactly like the reference's three
 * actors), or a single thread drives all three through the *_trio variants.
 *
 * Interfaces replaced (paths relative to the worldcoin/iris-mpc root):
 *   ShareDB            iris-mpc-gpu/src/dot/share_db.rs:448-1070
 *   ProcessedDatabase  iris-mpc-gpu/src/dot/share_db.rs:133-184
 *   Circuits           iris-mpc-gpu/src/threshold_ring/protocol.rs:433-2611
 *   DistanceComparator iris-mpc-gpu/src/dot/distance_comparator.rs:28-894
 *   NcclComm           iris-mpc-gpu/src/helpers/comm.rs:90-269
 *   derive_seed        iris-mpc-gpu/src/server/actor.rs:2914-2925
 */
#ifndef IRISMPC_H
#define IRISMPC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct irismpc_exchange irismpc_exchange;
typedef struct irismpc_share_db irismpc_share_db;
typedef struct irismpc_db irismpc_db;
typedef struct irismpc_query irismpc_query;
typedef struct irismpc_circuits irismpc_circuits;
typedef struct irismpc_comparator irismpc_comparator;
typedef struct irismpc_graph irismpc_graph;
typedef struct irismpc_batch_runner irismpc_batch_runner;
typedef struct irismpc_serving irismpc_serving;

/* ------------------------------------------------------------- misc */
/* Thread-local text of the last failure (empty string if none). */
const char* irismpc_last_error(void);
/* ABI version (major * 10000 + minor * 100 + patch). */
int irismpc_version(void);
int irismpc_device_count(int* n);
/* Streams and device memory for FFI callers without a HIP binding. */
int irismpc_stream_create(int device, void** stream);
int irismpc_stream_destroy(void* stream);
int irismpc_stream_synchronize(void* stream);
int irismpc_malloc(int device, size_t bytes, void** ptr);
int irismpc_free(void* ptr);
int irismpc_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int irismpc_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
/* Fills with a kernel (graph-capture safe). */
int irismpc_fill_u32(void* ptr, size_t n, uint32_t value, void* stream);
int irismpc_fill_u64(void* ptr, size_t n, uint64_t value, void* stream);
/* HKDF-SHA256 seed derivation of server/actor.rs:2914-2925 (KDF_SALT of
 * actor.rs:116, ikm = little-endian seed bytes, info = big-endian nonce). */
int irismpc_derive_seed(const uint32_t seed[8], uint64_t nonce, uint32_t out[8]);

/* ------------------------------------------------------------- exchange */
/* Replaces the per-GPU NcclComm of the three parties (helpers/comm.rs,
 * helpers/device_manager.rs:289-344). timeout_s bounds every rendezvous. */
int irismpc_exchange_create(int device, double timeout_s, irismpc_exchange** out);
int irismpc_exchange_destroy(irismpc_exchange* ex);

/* ------------------------------------------------------------- ShareDB */
/* ShareDB::init (share_db.rs:467-573). seeds = (own, prev) ChaCha12 seeds as
 * derived by derive_seed; comm == NULL builds a non-remote engine (no masking,
 * no reshare) exactly like an empty comms vector. tag distinguishes engines
 * sharing one exchange (codes, masks, ...). */
int irismpc_share_db_init(int device, int peer_id, size_t max_db_length, size_t query_length,
                          size_t code_length, const uint32_t seed_own[8],
                          const uint32_t seed_prev[8], irismpc_exchange* comm, int tag,
                          irismpc_share_db** out);
int irismpc_share_db_destroy(irismpc_share_db* e);
/* ShareDB::alloc_db (share_db.rs:575-618): the DB share stays resident in HBM
 * (the reference page-locks host memory and streams chunks). */
int irismpc_share_db_alloc_db(irismpc_share_db* e, size_t max_db_length, irismpc_db** out);
int irismpc_db_destroy(irismpc_db* db);
/* ProcessedDatabase::load_single_record_from_db for n consecutive records
 * (share_db.rs:186-195,620-655); records: n * code_length u16 on the host.
 * Per-record sums (preprocess, share_db.rs:265-296) are computed on load. */
int irismpc_db_load_records(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                            const uint16_t* records, void* stream);
/* load_single_record_from_s3 (share_db.rs:197-212,657-684): limb0 = low
 * ("odd") bytes, limb1 = high ("even") bytes, already biased by -128. */
int irismpc_db_load_records_limbs(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                                  const uint8_t* limb0, const uint8_t* limb1, void* stream);
/* InMemoryStore::load_single_record_from_s3 (inmemory_store.rs:55-105) for
 * a run of stored DB-chunk records in host memory, loaded straight into HBM.
 * Record layout (iris-mpc-store/src/s3_importer.rs:13-87, 76,806 bytes):
 * big-endian u32 serial id; left code, left mask, right code, right mask,
 * each as its low limb bytes ("odd") then its high limb bytes ("even"),
 * biased by -128; big-endian u16 version id. Serial id s goes to row s - 1
 * of the four slices (loader.rs:196-233); ids above max_serial_id (0 = no
 * cap) are skipped; id 0, an id repeated within the call or a row beyond a
 * slice fails before anything is written. *n_loaded = records loaded;
 * *row_end = largest loaded row + 1 (0 if none). */
int irismpc_db_load_stored_records(irismpc_share_db* codes, irismpc_share_db* masks,
                                   irismpc_db* left_code, irismpc_db* left_mask,
                                   irismpc_db* right_code, irismpc_db* right_mask,
                                   const uint8_t* records, size_t n_records,
                                   uint32_t max_serial_id, size_t* n_loaded, size_t* row_end,
                                   void* stream);
/* ProcessedDatabase::prefetch_subset (share_db.rs:344-407): DB rows ids[0..n)
 * (device u32) of db into rows 0..n of subset (a slice of at least n rows,
 * e.g. irismpc_share_db_alloc_db(e, 32768, ...)). Ids beyond the DB's
 * capacity leave their subset row as it was. */
int irismpc_db_prefetch_subset(irismpc_share_db* e, const irismpc_db* db, const uint32_t* ids,
                               size_t n, irismpc_db* subset, void* stream);
/* ProcessedDatabase::write_at_index (share_db.rs:214-262): copies the centre
 * rotation of query src_index into DB row dst_index. */
int irismpc_db_write_at_index(irismpc_share_db* e, irismpc_db* db, const irismpc_query* q,
                              size_t src_index, size_t dst_index, void* stream);
/* Synthetic DB content: uniformly random shares for rows [index, index + n)
 * (index a multiple of 16) from ChaCha20(key, nonce); sums are computed.
 * Uniform u16 shares have the distribution of real Shamir shares. */
int irismpc_db_fill_random(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                           const uint32_t key[8], uint64_t nonce, void* stream);
/* u16 rows already on the device (dev_rows, n x code_length) -> DB rows
 * [index, index + n); stream-ordered. No reference counterpart (the reference
 * loads from host memory only); used to load device-generated workloads. */
int irismpc_db_load_records_device(irismpc_share_db* e, irismpc_db* db, size_t index, size_t n,
                                   const uint16_t* dev_rows, void* stream);
/* Copies DB rows [src, src + n) to [dst, dst + n) on the device (multiples of
 * 16, non-overlapping). No reference counterpart: tiles a set of synthetic
 * irises across a large DB. */
int irismpc_db_replicate_records(irismpc_share_db* e, irismpc_db* db, size_t src, size_t n,
                                 size_t dst, void* stream);
/* Reads DB rows back as u16 (for tests). */
int irismpc_db_read_records(irismpc_share_db* e, const irismpc_db* db, size_t index, size_t n,
                            uint16_t* out, void* stream);
/* preprocess_query + htod_transfer_query + query_sums (share_db.rs:45-59,
 * 686-765; helpers/query_processor.rs:117-140): rows = query_length rows of
 * code_length u16 (rotations and Lagrange preprocessing already applied, as in
 * the reference's CompactQuery). */
int irismpc_query_upload(irismpc_share_db* e, const uint16_t* rows, size_t n_rows, void* stream,
                         irismpc_query** out);
/* Raw query shares of this party (batch x code_length u16, neither rotated nor
 * preprocessed, as received in the request): the Lagrange factor, basis B
 * (galois_engine.rs:23-40) and the 31 rotations (galois_engine.rs:42-52,
 * 479-488) are applied on the GPU; rows beyond batch*31 are zero padding up
 * to query_length (actor.rs:895-901 pads to max_batch_size). */
/* The same with exactly `rows` query rows (batch * 31 <= rows; rows may
 * exceed the engine's query_length): a packed query set of the batching form,
 * or a batch without padding. */
int irismpc_query_upload_shares_rows(irismpc_share_db* e, const uint16_t* shares, size_t batch,
                                     size_t rows, void* stream, irismpc_query** out);
/* Refills a query set made by irismpc_query_upload_shares[_rows] with new raw
 * shares (batch queries, `rows` rows, 0 = query_length), reusing its device
 * buffers: no allocation or free, so no implicit device-wide synchronisation.
 * rows and batch must not exceed what the set was made with. Replaces the
 * per-batch query buffer allocation of the reference's query processor
 * (iris-mpc-gpu/src/server/query_processor.rs:30-72) followed by
 * preprocess_query (dot/share_db.rs:45-59); used by a request loop whose
 * active query set changes between chunk steps. */
int irismpc_query_refill_shares(irismpc_share_db* e, irismpc_query* q, const uint16_t* shares,
                                size_t batch, size_t rows, void* stream);
/* irismpc_query_refill_shares without synchronising: the refill is ordered on
 * `stream` after the work already queued there (which may still read the
 * set) and before the work queued later, so a request loop changes its query
 * set between queued chunk steps without draining them. Every asynchronous
 * refill of one engine must use one stream (its staging buffers are reused in
 * stream order), and `shares` (pinned host memory) must stay unchanged until
 * the stream has run the copy. No reference counterpart. */
int irismpc_query_refill_shares_async(irismpc_share_db* e, irismpc_query* q, const uint16_t* shares,
                                      size_t batch, size_t rows, void* stream);
/* Queries of the Mirror orientation (full-face mirror attack pass,
 * actor.rs:591-615; server/mod.rs:156-176): the query rows come from the
 * mirrored shares of `shares` (the other eye's raw shares; mirrored_code for
 * 12,800-coefficient engines, mirrored_mask otherwise, as the server derives
 * them at iris-mpc/src/services/processors/mod.rs:39-40), their DB form
 * (batch self-comparison) from insert_shares (this eye's raw shares). */
int irismpc_query_upload_shares_mirrored(irismpc_share_db* e, const uint16_t* shares,
                                         const uint16_t* insert_shares, size_t batch,
                                         void* stream, irismpc_query** out);
int irismpc_query_upload_shares(irismpc_share_db* e, const uint16_t* shares, size_t batch,
                                void* stream, irismpc_query** out);
int irismpc_query_destroy(irismpc_query* q);
/* ShareDB::dot + ShareDB::dot_reduce_and_multiply (share_db.rs:800-905),
 * fused: DB rows [offset, offset + chunk_size) against all query rows; results
 * land in the engine's result buffer as [query_length][chunk_size] u16.
 * multiplier = 1 for codes, 2 for trimmed masks. */
int irismpc_share_db_dot_reduce(irismpc_share_db* e, const irismpc_query* q, const irismpc_db* db,
                                size_t chunk_size, size_t offset, uint16_t multiplier,
                                void* stream);
/* Consecutive irismpc_share_db_dot_reduce calls in one launch: DB rows
 * [offset, offset + total) as chunks of chunk_size rows (the last one
 * total mod chunk_size wide when that is not zero; total a multiple of 64),
 * results side by side in the engine's result buffer from element out_offset,
 * chunk c at out_offset + c * chunk_size * Qr, each with the layout and masks
 * of its own call (share_db.rs:800-905) made in order. No reference
 * counterpart: it serves the multi-chunk scan steps of a small query set. */
int irismpc_share_db_dot_reduce_chunks(irismpc_share_db* e, const irismpc_query* q,
                                       const irismpc_db* db, size_t offset, size_t total,
                                       size_t chunk_size, uint16_t multiplier,
                                       size_t out_offset, void* stream);
/* Whole-range form: rows [offset, offset+total) in chunks of chunk_size rows
 * in one launch; chunk c is written at out + c * chunk_size * Qr, Qr = the
 * query set's rows. No reference counterpart: one launch in place of the
 * per-chunk dot/dot_reduce pair (share_db.rs:800-905), with the masks at the
 * counters those calls would use. */
int irismpc_share_db_dot_reduce_range(irismpc_share_db* e, const irismpc_query* q,
                                      const irismpc_db* db, size_t offset, size_t total,
                                      size_t chunk_size, uint16_t multiplier, uint16_t* out,
                                      void* stream);
/* ShareDB::reshare_results (share_db.rs:1004-1036): collective. */
/* Batching form. n_batches query batches packed back to back in q (rows[i]
 * rows each, sum = the query set's rows; at most 64) in one launch: batch i's
 * results go to outs[i] with the layout of its own
 * irismpc_share_db_dot_reduce_range. Masks: with ctr_first == NULL the
 * counters those calls would use in order (the engine counter advances);
 * otherwise batch i starts at ctr_first[i] and the engine counter is left
 * alone. No reference counterpart (the reference runs one batch per launch). */
int irismpc_share_db_dot_reduce_packed(irismpc_share_db* e, const irismpc_query* q,
                                       const irismpc_db* db, size_t offset, size_t total,
                                       size_t chunk_size, uint16_t multiplier,
                                       const size_t* rows, uint16_t* const* outs, int n_batches,
                                       const uint64_t* ctr_first, void* stream);
int irismpc_share_db_reshare_results(irismpc_share_db* e, size_t chunk_size, void* stream);
int irismpc_share_db_reshare_results_trio(irismpc_share_db* const e[3], size_t chunk_size,
                                          void* const streams[3]);
/* ShareDB::result_chunk_shares (share_db.rs:1050-1069): (a, b) = (results,
 * results_peer) device pointers. On one GPU the reshare hands b over as a
 * view of the previous party's own result buffer (no copy): it stays valid
 * until that party's next dot_reduce. Engines that need b to outlive it set
 * irismpc_share_db_set_peer_copy. */
int irismpc_share_db_result_views(irismpc_share_db* e, const uint16_t** a, const uint16_t** b);
/* 1: the reshare copies the previous party's share into this engine's own
 * buffer (when any of the three engines asks for it); 0 (default): view.
 * No reference counterpart (the reference's NCCL receive always copies). */
int irismpc_share_db_set_peer_copy(irismpc_share_db* e, int copy);
/* ShareDB::fetch_results (share_db.rs:1038-1048). */
int irismpc_share_db_fetch_results(irismpc_share_db* e, uint16_t* out, size_t n, void* stream);
int irismpc_share_db_get_counter(const irismpc_share_db* e, uint64_t* counter);
int irismpc_share_db_set_counter(irismpc_share_db* e, uint64_t counter);
/* GEMM kernel variant: -1 (default) = automatic: the rotation-aware kernel for
 * query sets uploaded from raw shares (irismpc_query_upload_shares; its
 * two-query tile for sets of 1, 2, 5, 6, 9 or 10 queries, where the scan is
 * bound by DB reads and the four-query tile's last tile would be at most half
 * full, its four-query tile otherwise), the
 * best generic tile otherwise; 0 .. irismpc_gemm_variants()-1 = generic tiles;
 * irismpc_gemm_rot_variants() = rotation-aware tiles (raw-share queries only).
 * No reference counterpart: the reference has one cuBLAS-backed GEMM path. */
int irismpc_share_db_set_variant(irismpc_share_db* e, int variant);
/* Launch timing, off by default. No reference counterpart: with on != 0,
 * every share GEMM launch of the engine is bracketed by HIP events recorded on
 * its stream (nothing synchronises at launch); launch_times waits for the
 * recorded events and returns the number of timed launches and their summed
 * durations (ms) since its previous call. */
int irismpc_share_db_set_timing(irismpc_share_db* e, int on);
int irismpc_share_db_launch_times(irismpc_share_db* e, size_t* launches, double* total_ms);
int irismpc_gemm_variants(int* n);
int irismpc_gemm_rot_variants(int* first, int* n);

/* ------------------------------------------------------------- Circuits */
/* Circuits::new (protocol.rs:463-508): input_size elements per call,
 * alloc_size words (elements/64) of scratch. */
int irismpc_circuits_new(int device, int peer_id, size_t input_size, size_t alloc_size,
                         const uint32_t seed_own[8], const uint32_t seed_prev[8],
                         irismpc_exchange* comm, int tag, irismpc_circuits** out);
int irismpc_circuits_destroy(irismpc_circuits* c);
/* Circuits::set_chunk_size (protocol.rs:510-515), in 64-element words. */
int irismpc_circuits_set_chunk_size(irismpc_circuits* c, size_t chunk_words);
/* Circuits::compare_threshold_masked_many (protocol.rs:2519-2546): code/mask
 * replicated u16 shares (a, b) of len elements (multiple of 512); result bit
 * shares in plane 0 of the result buffer. Collective. */
int irismpc_circuits_compare_threshold_masked_many(irismpc_circuits* c, const uint16_t* code_a,
                                                   const uint16_t* code_b,
                                                   const uint16_t* mask_a,
                                                   const uint16_t* mask_b, size_t len,
                                                   uint32_t threshold_a, void* stream);
int irismpc_circuits_compare_threshold_masked_many_trio(
    irismpc_circuits* const c[3], const uint16_t* const code_a[3], const uint16_t* const code_b[3],
    const uint16_t* const mask_a[3], const uint16_t* const mask_b[3], size_t len,
    uint32_t threshold_a, void* const streams[3]);
/* take_result_buffer (protocol.rs:520-526): device pointers of the result
 * shares (a, b), chunk_size words each. */
int irismpc_circuits_result_buffer(irismpc_circuits* c, const uint64_t** a, const uint64_t** b);
/* The exchange inside actor.rs:2928-2997 `open`: own b share to next, prev's b
 * into this party's c view (the previous party's buffer itself, valid until
 * the next compare). Collective. */
int irismpc_circuits_open(irismpc_circuits* c, void* stream);
int irismpc_circuits_open_trio(irismpc_circuits* const c[3], void* const streams[3]);
int irismpc_circuits_open_views(irismpc_circuits* c, const uint64_t** a, const uint64_t** b,
                                const uint64_t** cc);
int irismpc_circuits_get_counter(const irismpc_circuits* c, uint64_t* counter);
int irismpc_circuits_set_counter(irismpc_circuits* c, uint64_t counter);
/* Launch timing of the circuits, as irismpc_share_db_set_timing: one timed
 * interval per threshold comparison (lift, bit injection and MSB adder of the
 * three parties on the protocol stream), kept by party 0's engine. */
int irismpc_circuits_set_timing(irismpc_circuits* c, int on);
int irismpc_circuits_launch_times(irismpc_circuits* c, size_t* launches, double* total_ms);
/* Debug capture of the corrected lifted mask (u32 a, b), NULL to disable. */
int irismpc_circuits_set_lift_capture(irismpc_circuits* c, uint32_t* a, uint32_t* b);

/* ------------------------------------------------------------- DistanceComparator */
/* DistanceComparator::init (distance_comparator.rs:58-185). */
int irismpc_comparator_init(int device, size_t query_length, size_t max_db_size,
                            irismpc_comparator** out);
int irismpc_comparator_destroy(irismpc_comparator* d);
/* DistanceComparator::open_results (distance_comparator.rs:196-278) without
 * the anon-stats capture: bit = a ^ b ^ c, OR over rotations into
 * bitmap[total_db_len * (q / 31) + db + offset]. */
int irismpc_comparator_open_results(irismpc_comparator* d, const uint64_t* a, const uint64_t* b,
                                    const uint64_t* c, uint64_t* bitmap, size_t chunk_len,
                                    size_t real_chunk_len, size_t offset, size_t total_db_len,
                                    size_t batch_size, void* stream);
/* join_db_matches / join_db_matches_with_bitmaps (distance_comparator.rs:
 * 487-543,713-793): AND of the eyes, or per-pair OR where or_policy has the
 * bit set (or_policy == NULL: AND everywhere). final_results: u32 per query,
 * initialised to UINT32_MAX by the caller. */
int irismpc_comparator_join_db_matches(irismpc_comparator* d, const uint64_t* left,
                                       const uint64_t* right, uint32_t* final_results,
                                       size_t db_size, const uint64_t* or_policy,
                                       size_t max_db_size, void* stream);
/* DistanceComparator::get_partial_results (distance_comparator.rs:545-648)
 * with the OR-rule / reauth indices the actor appends (actor.rs:946-970):
 * the ascending list, without repeats, of DB ids d < db_size set in any of
 * the batch_size bitmap rows (row_bits bits apart; 0 = db_size) or listed in
 * extra_ids (host, n_extra; ids >= db_size are skipped). The list stays on the
 * device (irismpc_comparator_partial_ids); *n_ids receives its length and,
 * when host_ids is not NULL, its first host_capacity ids are copied there.
 * The reference keeps at most DB_CHUNK_SIZE set bits (dropping the rest);
 * this list keeps every id. Synchronises the stream. */
int irismpc_comparator_get_partial_results(irismpc_comparator* d, const uint64_t* bitmap,
                                           size_t db_size, size_t batch_size, size_t row_bits,
                                           const uint32_t* extra_ids, size_t n_extra,
                                           uint32_t* host_ids, size_t host_capacity,
                                           size_t* n_ids, void* stream);
/* DistanceComparator::join_batch_matches (distance_comparator.rs:731-746,
 * mergeBatchResults dot/kernel.cu:352-386) over the batch self-comparison
 * bitmaps of irismpc_scan_batch_self_trio (bit row_bits * q + j: query q
 * matches query j; row_bits 0 = batch_size): for j < q matching in both
 * eyes, final_results[q] = min(final_results[q], 0xFFFFFFFE) and
 * 0xFFFFFFFF - j is appended to q's match list (counters as for DB matches). */
int irismpc_comparator_join_batch_matches(irismpc_comparator* d, const uint64_t* left,
                                          const uint64_t* right, uint32_t* final_results,
                                          size_t batch_size, size_t row_bits, void* stream);
/* Match-distance capture for anonymized statistics (openResults,
 * dot/kernel.cu:54-116; storeAnonStatsWithIndexMapping :153-210; buffers of
 * server/anon_stats.rs:13-123) in the opens that follow: for every opened
 * match (and every rotation of each target pair db * (max_query_length / 31)
 * + query from targets, host u64, matching or not) the four dot-product
 * shares (code a/b, mask a/b, u16) and the match id batch_id * max_db_length
 * * max_query_length + db * max_query_length + query row (u64) are appended
 * to the device buffers (capacity entries; *counter counts every entry).
 * irismpc_scan_db* capture in their opens; irismpc_scan_subset_trio captures
 * in its separate anon_threshold pass. capacity 0 switches the capture off. */
int irismpc_comparator_set_anon_capture(irismpc_comparator* d, uint16_t* code_a, uint16_t* code_b,
                                        uint16_t* mask_a, uint16_t* mask_b, uint64_t* idx,
                                        uint32_t* counter, size_t capacity, uint64_t batch_id,
                                        size_t max_query_length, size_t max_db_length,
                                        const uint64_t* targets, size_t n_targets);
/* The candidate list of irismpc_comparator_get_partial_results computed
 * without allocating device memory and without synchronising: no counterpart
 * in the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops). On `stream`: the ascending list, without
 * repeats, of DB ids d < db_size set in any of the batch_size bitmap rows
 * (row_bits bits apart; 0 = db_size) or listed in extra_ids (device u32,
 * n_extra; ids >= db_size skipped) goes to ids (device, room for db_size
 * ids) and its length to *count (device u32). Its scratch is a device array
 * of the library, 16 MiB per GPU for db_size up to 2^27 rows (134,217,728);
 * a larger db_size fails before anything runs. Calls on one GPU share that
 * array: calls that overlap in time on one GPU race. */
int irismpc_partial_results_noalloc(const uint64_t* bitmap, size_t db_size, size_t batch_size,
                                    size_t row_bits, const uint32_t* extra_ids, size_t n_extra,
                                    uint32_t* ids, uint32_t* count, void* stream);
/* Device pointer of the last get_partial_results list. */
int irismpc_comparator_partial_ids(irismpc_comparator* d, const uint32_t** ids);
/* open_results_with_index_mapping (distance_comparator.rs:281-342,
 * dot/kernel.cu:118-151): column j < real_chunk_len of the opened subset
 * chunk is DB row index_map[j] (device u32); each match sets bitmap bit
 * total_db_len * (q / 31) + index_map[j] and is captured as (q / 31,
 * index_map[j], q % 31 - 15). */
int irismpc_comparator_open_results_with_index_mapping(
    irismpc_comparator* d, const uint64_t* a, const uint64_t* b, const uint64_t* c,
    uint64_t* bitmap, size_t chunk_len, size_t real_chunk_len, size_t total_db_len,
    size_t batch_size, const uint32_t* index_map, void* stream);
/* get_partial_results_with_rotations (distance_comparator.rs:650-711): copies
 * up to capacity captured (query, db id, rotation) triples to the host arrays
 * and resets the capture; *n_matches = matches seen (the capture holds the
 * first 256 * query_length). */
int irismpc_comparator_fetch_partial_rotations(irismpc_comparator* d, uint32_t* query,
                                               uint32_t* db, int8_t* rotation, size_t capacity,
                                               size_t* n_matches, void* stream);
/* Resets match counters (start of a batch). */
int irismpc_comparator_reset(irismpc_comparator* d, void* stream);
/* Device pointers of match_counters, all_matches (256 per query),
 * match_counters_left/right, partial_results_left/right. */
int irismpc_comparator_buffers(irismpc_comparator* d, uint32_t** match_counters,
                               uint32_t** all_matches, uint32_t** counters_left,
                               uint32_t** counters_right, uint32_t** partial_left,
                               uint32_t** partial_right);

/* ------------------------------------------------------------- DB scan */
/* One party's full-DB scan of a query batch: the chunk loop of
 * ServerActor::compare_query_against_db_and_self_prefiltering (actor.rs:
 * 2246-2589) - per 32,768-row chunk: dot + dot_reduce (codes, masks),
 * reshare_results x2, compare_threshold_masked_many(threshold_a), open,
 * open_results into bitmap. Collective: the three parties call concurrently. */
int irismpc_scan_db(irismpc_share_db* codes, irismpc_share_db* masks, irismpc_circuits* circuits,
                    irismpc_comparator* cmp, const irismpc_db* code_db, const irismpc_db* mask_db,
                    const irismpc_query* q_code, const irismpc_query* q_mask, size_t db_size,
                    size_t batch_size, uint32_t threshold_a, uint64_t* bitmap, void* stream);
/* The three parties' scans from one call. Without a distance capture the loop
 * runs on the calling thread: each chunk's six share GEMMs (three parties x
 * code, mask) go in order on streams[0] and the circuits and opens on the
 * exchange's protocol stream, overlapping the next chunk's GEMMs; with a
 * capture, one host thread per party runs irismpc_scan_db. Same shares,
 * counters and bitmaps either way. */
int irismpc_scan_db_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                         irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                         const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                         const irismpc_query* const q_code[3], const irismpc_query* const q_mask[3],
                         size_t db_size, size_t batch_size, uint32_t threshold_a,
                         uint64_t* const bitmaps[3], void* const streams[3]);
/* One chunk step of those scans (the body of the chunk loop of actor.rs:
 * 2318-2533) for the three parties: DB rows [offset, offset + min(32768,
 * db_size - offset)) against the query sets (their own rows: batch_size
 * queries x 31 or fewer padding rows). slot_of[p] (device u32 per batch
 * query, or NULL) sends query j's bits to bitmap row slot_of[p][j]; rows are
 * row_bits bits apart (0: db_size, the layout of irismpc_scan_db; a nonzero
 * row_bits below db_size is rejected before any launch). The
 * serving loop's iteration: requests join and leave the query set between
 * chunk steps. */
int irismpc_scan_chunk_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                            irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                            const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                            const irismpc_query* const q_code[3], const irismpc_query* const q_mask[3],
                            size_t offset, size_t db_size, size_t batch_size, uint32_t threshold_a,
                            uint64_t* const bitmaps[3], const uint32_t* const slot_of[3],
                            size_t row_bits, void* const streams[3]);
/* n_chunks (1 .. 16) consecutive chunk steps from `offset` in one step: each
 * chunk's dot shares sit side by side in the engines' result buffers
 * (n_chunks x 32,768 x rows must fit them; one GEMM launch per matrix covers
 * them, as irismpc_share_db_dot_reduce_chunks), one reshare and one threshold
 * comparison cover them, and each chunk opens into the bitmaps as above. The
 * opened bits equal those of n_chunks single steps; the circuit randomness
 * follows one call's schedule. No reference counterpart: the reference's
 * loop body is one chunk; a request loop with few active queries uses this to
 * fill the GPU with fewer, larger launches. */
int irismpc_scan_chunks_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                             irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                             const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                             const irismpc_query* const q_code[3], const irismpc_query* const q_mask[3],
                             size_t offset, size_t n_chunks, size_t db_size, size_t batch_size,
                             uint32_t threshold_a, uint64_t* const bitmaps[3],
                             const uint32_t* const slot_of[3], size_t row_bits, void* const streams[3]);
/* The DB half of ServerActor::compare_query_against_db_subset_and_self
 * (actor.rs:1900-2244) for the three parties: the subset ids[p][0..n_ids)
 * (device u32 DB row ids, e.g. irismpc_comparator_partial_ids) are gathered
 * from code_db/mask_db into code_sub/mask_sub (slices of DB_CHUNK_SIZE rows)
 * DB_CHUNK_SIZE at a time, compared at threshold_a, opened with the index
 * mapping into bitmaps[p] (rows total_db_len bits apart) and captured with
 * their rotations. anon_threshold != 0: every chunk is first compared at it
 * for the distance capture alone (the db_threshold_anon_stats pass,
 * actor.rs:2040-2120), when the comparator has a capture set. */
int irismpc_scan_subset_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                             irismpc_circuits* const circuits[3], irismpc_comparator* const cmp[3],
                             const irismpc_db* const code_db[3], const irismpc_db* const mask_db[3],
                             irismpc_db* const code_sub[3], irismpc_db* const mask_sub[3],
                             const irismpc_query* const q_code[3],
                             const irismpc_query* const q_mask[3], const uint32_t* const ids[3],
                             size_t n_ids, size_t total_db_len, size_t batch_size,
                             uint32_t threshold_a, uint32_t anon_threshold,
                             uint64_t* const bitmaps[3], void* const streams[3]);
/* ServerActor::compare_query_against_self (actor.rs:1789-1897) for the three
 * parties: the batch's queries (all rotations) against the batch's own
 * irises in DB form at threshold_a; bit row_bits * q + j of bitmaps[p] = some
 * rotation of query q matches query j (row_bits 0 = batch_size). Needs query
 * sets uploaded from raw shares (irismpc_query_upload_shares[_rows]). The
 * reference also computes the rotated DB-side columns, which its merge never
 * reads; they are skipped here. */
int irismpc_scan_batch_self_trio(irismpc_share_db* const codes[3], irismpc_share_db* const masks[3],
                                 irismpc_circuits* const circuits[3],
                                 irismpc_comparator* const cmp[3],
                                 const irismpc_query* const q_code[3],
                                 const irismpc_query* const q_mask[3], size_t batch_size,
                                 uint32_t threshold_a, uint64_t* const bitmaps[3], size_t row_bits,
                                 void* const streams[3]);

/* ------------------------------------------------ batching form */
/* Query batches of uneven size packed into few launches. The reference runs
 * batch after batch, every query set padded to max_batch_size x 31 rows
 * (iris-mpc-gpu/src/server/actor.rs:895-901, share_db.rs:800-845); the
 * packing, its planner and its schedule have no reference counterpart. The
 * three host-only functions below need no GPU. */
/* Greedy packing of consecutive batches (sizes[i] queries each): a pack holds
 * at most max_queries queries (a larger batch gets one of its own) and at most
 * max_batches batches. pack_of_batch[i] = pack of batch i; *n_packs = packs. */
int irismpc_batch_plan_packs(const uint32_t* sizes, size_t n_batches, size_t max_queries,
                             size_t max_batches, uint32_t* pack_of_batch, size_t* n_packs);
/* ChaCha blocks one remote dot_reduce_range over `total` rows uses (mask and
 * reshare OTP per chunk), and a whole scan of n_rows rows launched chunk by
 * chunk with the last chunk padded to 64 rows. */
int irismpc_batch_counter_advance(size_t total, size_t chunk, size_t rows, uint64_t* blocks);
int irismpc_batch_scan_counter_advance(size_t n_rows, size_t chunk, size_t rows, uint64_t* blocks);
/* One matrix (code or mask) of one party: batches of sizes[i] queries whose
 * raw shares sit back to back in `shares` (sum(sizes) x code_length u16, host);
 * outs[i] (device) receives batch i's results as its own dot_reduce_range over
 * the DB's n_rows rows lays them out (the last chunk stays: chunk x sizes[i] x
 * 31 u16 + 64). The engine's counter is the plain form's counter. */
int irismpc_batch_runner_create(irismpc_share_db* e, const irismpc_db* db, size_t n_rows,
                                uint16_t multiplier, const uint32_t* sizes, size_t n_batches,
                                const uint16_t* shares, uint16_t* const* outs, size_t chunk,
                                size_t max_queries, size_t max_batches, void* stream,
                                irismpc_batch_runner** out);
int irismpc_batch_runner_destroy(irismpc_batch_runner* r);
/* New pack size (rebuilds the plan and the packed query sets). */
int irismpc_batch_runner_replan(irismpc_batch_runner* r, size_t max_queries, void* stream);
int irismpc_batch_runner_plan(const irismpc_batch_runner* r, uint32_t* pack_of_batch, size_t* n_packs);
/* The packed schedule over the first n_rows rows (0 = all): one launch per
 * pack and chunk, each batch at the counters its plain run from ctr0 would
 * use. Allocates and synchronises nothing: it can be captured into a graph. */
int irismpc_batch_runner_issue_packed(irismpc_batch_runner* r, uint64_t ctr0, size_t n_rows,
                                      void* stream);
/* Batch `batch`'s plain run over the whole DB at the engine's counter. */
int irismpc_batch_runner_issue_plain(irismpc_batch_runner* r, size_t batch, void* stream);
/* Persistent form of the packed schedule: one launch per chunk covering every
 * pack, a block per CU looping over the chunk's tiles (XCD-aware), with the
 * outputs and counters of issue_packed(ctr0) over the whole DB. prepare builds
 * the launch descriptors in device memory (allocates, synchronises); issue
 * allocates and synchronises nothing and can be captured into a graph. At most
 * 64 packs; every pack runs the 4-query rotation-aware tile. */
int irismpc_batch_runner_prepare_persistent(irismpc_batch_runner* r, uint64_t ctr0, void* stream);
int irismpc_batch_runner_issue_persistent(irismpc_batch_runner* r, void* stream);

/* ------------------------------------------------ serving form */
/* Request loops over the three parties' HBM-resident DB shares, with their own
 * engines: the continuous policy (uniqueness requests join the active query
 * set between chunk steps as slots and free HBM allow; no reference
 * counterpart) and the per-batch policy (ServerActor's batch after batch,
 * actor.rs:584-644, each batch padded to max_batch x 31 query rows, one full
 * scan per batch). Reauth requests run on a one-query engine set over the
 * 64-row window of their target; insertions append a row, deletions overwrite
 * one with shares of the default iris (get_dummy_shares_for_deletion,
 * iris-mpc-common/src/iris_db/mod.rs:10-23). */
#define IRISMPC_REQ_UNIQUENESS 0
#define IRISMPC_REQ_REAUTH 1
#define IRISMPC_REQ_INSERT 2
#define IRISMPC_REQ_DELETE 3
#define IRISMPC_SERVING_CONTINUOUS 0
#define IRISMPC_SERVING_PER_BATCH 1
#define IRISMPC_REQ_PENDING 0
#define IRISMPC_REQ_DONE 1
#define IRISMPC_REQ_FAILED (-1)
/* party_seeds: [3][8] u32, party p's seed (its engines' prev seed is party
 * p - 1's). The DB slices (any engine's alloc_db of >= capacity rows, matching
 * code lengths) hold n_rows rows and stay owned by the caller; capacity bounds
 * insertions (capacity - 32,768 rows). max_active: 1 .. 64 slots. */
int irismpc_serving_create(int device, const uint32_t* party_seeds, irismpc_db* const code_dbs[3],
                           irismpc_db* const mask_dbs[3], size_t n_rows, size_t capacity,
                           size_t max_active, uint32_t threshold_a, irismpc_serving** out);
int irismpc_serving_destroy(irismpc_serving* s);
/* Starts the loop thread and the reauth worker. policy: IRISMPC_SERVING_*;
 * max_batch: the per-batch policy's padding (1 .. max_active); reserve_bytes:
 * free HBM that admission keeps. */
int irismpc_serving_start(irismpc_serving* s, int policy, size_t max_batch, size_t reserve_bytes);
/* Waits until every submitted request has finished (timeout_s <= 0: no
 * limit), then stops the loop; fails if the loop failed or the wait timed out. */
int irismpc_serving_finish(irismpc_serving* s, double timeout_s);
/* code_shares [3][12800] u16 and mask_shares [3][6400] u16: the parties' raw
 * shares (null for deletions); target: reauth / deletion row; arrival_s: the
 * request's arrival on irismpc_serving_now's clock (< 0: now); a later arrival
 * is held until its time, so a whole schedule can be submitted at once. */
int irismpc_serving_submit(irismpc_serving* s, int kind, const uint16_t* code_shares,
                           const uint16_t* mask_shares, int64_t target, double arrival_s,
                           uint64_t* id);
/* state: IRISMPC_REQ_*; value: uniqueness match count, reauth 0/1, insertion /
 * deletion row; matches (uniqueness, ascending) copied up to max_matches,
 * *n_matches = their number; times[0] arrival, times[1] completion (s);
 * db_rows: the DB rows the request was checked against; error: the failure
 * text (error_len bytes incl. the terminator). Pointers may be null. */
int irismpc_serving_result(const irismpc_serving* s, uint64_t id, int* state, int64_t* value,
                           uint32_t* matches, size_t max_matches, size_t* n_matches, double* times,
                           size_t* db_rows, char* error, size_t error_len);
/* The DB version of a finished request's answer: the number of the loop's
 * DB writes (insertions and deletions that succeeded, counted from 1 in the
 * order the loop applied them) that the answer reflects. A uniqueness answer
 * is that of the DB after exactly these writes, whichever of them happened
 * while the request was scanning (the serving form scans a written chunk
 * again for every active request); a write reports its own position. No
 * counterpart in the reference, whose batches see the writes applied before
 * them (actor.rs:584-644); it states the same order for the serving form. */
int irismpc_serving_result_version(const irismpc_serving* s, uint64_t id, uint64_t* db_version);
/* The in-flight matches of a finished uniqueness request: the ids of the
 * uniqueness requests admitted before it and still scanning at its admission
 * (the same batch, before it, in the per-batch form) whose iris it matches at
 * the loop's threshold, over its 31 rotations, as the reference marks a
 * batch's later query that matches an earlier one (compare_query_against_self
 * and mergeBatchResults, actor.rs:1789-1897, dot/kernel.cu:352-386, whose
 * final result UINT_MAX - 1 flags such a query). Ids copied up to max_ids in
 * admission order, *n_ids = their number. */
int irismpc_serving_result_batch_matches(const irismpc_serving* s, uint64_t id, uint64_t* ids, size_t max_ids,
                                         size_t* n_ids);
/* Forgets a finished request (fails while it is pending). A finished
 * request's shares are dropped at once; its result stays until released, so a
 * long-running host releases every request it has read. */
int irismpc_serving_release(irismpc_serving* s, uint64_t id);
int irismpc_serving_now(const irismpc_serving* s, double* seconds);
/* out[5]: chunk steps, largest active set, admissions deferred by the reserve,
 * per-batch batches, query-set refills (since start). */
int irismpc_serving_stats(const irismpc_serving* s, uint64_t* out);
/* out[5]: host seconds of the loop thread since start in query-set refills,
 * waits for the queued steps (at most two stay queued), queueing steps,
 * free-HBM checks and bitmap decoding. */
int irismpc_serving_loop_times(const irismpc_serving* s, double* out);
int irismpc_serving_n_rows(const irismpc_serving* s, size_t* n);
/* Loop stopped: one uniqueness query alone over the whole DB (the plain
 * per-request form); matches as in irismpc_serving_result. */
int irismpc_serving_run_alone(irismpc_serving* s, const uint16_t* code_shares,
                              const uint16_t* mask_shares, uint32_t* matches, size_t max_matches,
                              size_t* n_matches);
/* Loop stopped: seconds of one scan step of `batch` queries over n_chunks
 * chunks (best of reps after a warm-up). */
int irismpc_serving_time_step(irismpc_serving* s, size_t batch, size_t n_chunks, int reps,
                              double* seconds);
/* Loop stopped: admission of the continuous policy. Requests join while fewer
 * than active_cap are active, and every waiting request joins (slots and HBM
 * allowing) while more than queue_threshold wait. Default: active_cap =
 * max_active (every request a slot allows). A request's latency is the DB's
 * chunks times the step time of the active set, so a small cap keeps it short;
 * the threshold buys throughput back when a queue builds up. */
int irismpc_serving_set_admission(irismpc_serving* s, size_t active_cap, size_t queue_threshold);
/* Loop stopped: chunks per step for active counts from buckets[i] up to the
 * next bucket (n = 0: the fixed rule 8 / 4 / 2 / 1 chunks for 1 / 2 / <= 8 /
 * more active queries). */
int irismpc_serving_set_step_table(irismpc_serving* s, const uint32_t* buckets,
                                   const uint32_t* chunks, size_t n);
/* Loop stopped: times one step for every bucket (<= max_active) and every k of
 * ks that the result buffers and the DB allow, picks per bucket with
 * irismpc_serving_pick_step and sets the table. chunks_out[n_buckets] (0 past
 * max_active); seconds_per_chunk_out[n_buckets][n_ks] (negative: not timed). */
int irismpc_serving_tune_steps(irismpc_serving* s, const uint32_t* buckets, size_t n_buckets,
                               const uint32_t* ks, size_t n_ks, double tol, double max_step_s,
                               uint32_t* chunks_out, double* seconds_per_chunk_out);
int irismpc_serving_chunks_per_step(const irismpc_serving* s, size_t batch, size_t chunk_index,
                                    size_t n_chunks, size_t* k);
/* Host only: a step size from measured seconds per chunk: steps longer than
 * max_step_s (<= 0: no limit) are left out (all longer: the smallest k); of
 * the rest the smallest k within tol of the fastest per chunk. */
int irismpc_serving_pick_step(const uint32_t* ks, const double* seconds_per_chunk, size_t n,
                              double tol, double max_step_s, uint32_t* k);

/* ------------------------------------------------ HIP graphs */
/* Captures the library calls issued on `stream` between begin and end into a
 * graph (hipStreamCaptureModeThreadLocal) and instantiates it; launch replays
 * it. Captured calls must not allocate or synchronize (share GEMM launches,
 * packed launches and circuit steps qualify; the library clears memory with
 * fill kernels, never hipMemsetAsync). Engine counters advance once, at
 * capture: every replay reuses the captured masks. No reference counterpart. */
int irismpc_graph_capture_begin(void* stream);
int irismpc_graph_capture_end(void* stream, irismpc_graph** out);
int irismpc_graph_launch(irismpc_graph* g, void* stream);
int irismpc_graph_destroy(irismpc_graph* g);

/* ------------------------------------------------ synthetic workloads */
/* No reference counterpart: the data generator behind the benchmarks and the
 * full-size tests (the role of iris_db::db::IrisDB::new_random_rng and
 * GaloisRingSharedIris::generate_shares_locally in the reference's benches).
 * ChaCha20 streams under the all-zero key; see csrc/synth.hip for the stream
 * definitions. codes/masks are [n][200] u64 (bit i of word i / 64). */
/* Host: plaintext iris (dataset, index). */
int irismpc_synth_iris(uint64_t dataset, uint64_t index, uint64_t code[200], uint64_t mask[200]);
/* Host: flips n_flips distinct code bits in place (draws from (dataset, index)). */
int irismpc_synth_flip(uint64_t code[200], uint64_t dataset, uint64_t index, uint32_t n_flips);
/* Device: plaintext irises index0 .. index0 + n - 1 into device arrays. */
int irismpc_synth_plain_device(uint64_t dataset, uint64_t index0, size_t n, uint64_t* codes,
                               uint64_t* masks, void* stream);
/* Device: the three parties' shares of n device-resident plaintext irises
 * (iris i shares with randomness (rand_dataset, rand_index0 + i)): code_out[p]
 * [n][12800] u16, mask_out[p] [n][6400] u16 (trimmed mask). */
int irismpc_synth_shares_device(const uint64_t* codes, const uint64_t* masks, size_t n,
                                uint64_t rand_dataset, uint64_t rand_index0,
                                uint16_t* const code_out[3], uint16_t* const mask_out[3],
                                void* stream);

#ifdef __cplusplus
}
#endif
#endif
