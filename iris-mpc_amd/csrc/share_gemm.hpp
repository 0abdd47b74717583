// Share GEMM: exact Z_2^16 dot products of u16 shares via int8 MFMA.
#pragma once
#include "common.hpp"

namespace irismpc {

// Arguments of the fused GEMM + correction + zero-share-mask kernel.
// Computes, for query rows q < qr and DB rows r in [db_begin, db_end):
//   c = A0.B0 + 256 (A0.B1 + A1.B0)             (share_db.rs:800-845)
//   v = u16(c + (s0 << 7) + ((s0 + s1) << 15))  (dot/kernel.cu:16-28)
//   out = v * mult + F(own)[idx] - F(prev)[idx]  (masks only if masked != 0)
// with the reference's chunk geometry, relative to db_begin: rows are grouped
// into chunks of `chunk` rows, chunk c has width
// w_c = min(chunk, (db_end - db_begin) - c*chunk) and the result element index
// inside the chunk is idx = q * w_c + (r - db_begin - c*chunk).
// A packed launch carries several query batches back to back. Each keeps the
// output array, result layout (idx = q * w_c + local with its own q) and mask
// counters of its own plain launch, so the shares are those of per-batch runs.
constexpr int kMaxSegments = 64;
struct GemmSegment {
  uint16_t* out;            // chunk c at out + c * out_chunk_stride
  size_t out_chunk_stride;  // chunk * rows
  uint64_t ctr_first;       // ChaCha block counter of the batch's chunk 0
  uint64_t ctr_per_chunk;
  int row_begin;            // first row of the batch in the packed query set
  int rows;
};

// One launch's arguments without the packed-batch table.
struct GemmHead {
  static constexpr bool kPacked = false;
  const int8_t* q;      // tiled query limbs (rows padded to the q tile)
  const int8_t* db;     // tiled DB limbs
  const int8_t* qrot;   // padded unrotated query limb planes (rotation-aware variants)
  const uint32_t* qsum0;
  const uint32_t* qsum1;
  const uint32_t* dsum0;
  const uint32_t* dsum1;
  uint16_t* out;        // results, chunk c at out + c*out_chunk_stride
  size_t out_chunk_stride;
  int n_kt;             // K / 64
  int qr;               // query rows produced (query_length)
  int n_qtiles;
  int n_dtiles;
  size_t db_begin;      // first DB row (multiple of 16)
  size_t db_end;        // padded end (db_end - db_begin multiple of 64)
  size_t chunk;         // rows per chunk (multiple of 256)
  uint64_t ctr_first;   // ChaCha block counter of chunk 0's mask fill
  uint64_t ctr_per_chunk;
  uint64_t nonce;
  ChaChaKey key_own, key_prev;
  uint32_t mult;
  int masked;
};

struct GemmArgs : GemmHead {
  static constexpr bool kPacked = true;
  int n_seg;                // 0: one batch (out, ctr_first, ...); else the table below
  GemmSegment seg[kMaxSegments];
};

// Tiles u16 rows [n_rows][K] into the fragment layout at row offset row0 and
// writes per-row limb sums (share_db.rs:45-59,265-296).
void launch_tile_rows(const uint16_t* src, size_t n_rows, int K, int8_t* dst, size_t row0,
                      uint32_t* sum0, uint32_t* sum1, hipStream_t s);
// Same from split limbs (S3 even/odd format, share_db.rs:657-684).
void launch_tile_limbs(const uint8_t* l0, const uint8_t* l1, size_t n_rows, int K, int8_t* dst,
                       size_t row0, uint32_t* sum0, uint32_t* sum1, hipStream_t s);
// Stored records (iris-mpc-store/src/s3_importer.rs:13-87): n records of
// `stride` bytes on the device, each opening with its big-endian u32 serial
// id; the matrix's low-limb bytes start at off_lo and its high-limb bytes at
// off_hi within the record. Record with id in [1, max_serial] -> row id - 1.
void launch_tile_records(const uint8_t* records, size_t n, size_t stride, size_t off_lo,
                         size_t off_hi, int K, uint32_t max_serial, int8_t* dst, uint32_t* sum0,
                         uint32_t* sum1, hipStream_t s);
// Copies tiled row `src_row` of `src` to row `dst_row` of `dst` with its sums
// (write_at_index, share_db.rs:187-240).
void launch_copy_row(const int8_t* src, const uint32_t* src_s0, const uint32_t* src_s1,
                     size_t src_row, int8_t* dst, uint32_t* dst_s0, uint32_t* dst_s1,
                     size_t dst_row, int K, hipStream_t s);
// prefetch_subset (share_db.rs:344-407): tiled rows ids[0..n) of `src` (ids
// below src_rows; others leave their row alone) to rows 0..n of `dst`, sums
// included.
void launch_gather_rows(const int8_t* src, const uint32_t* src_s0, const uint32_t* src_s1,
                        size_t src_rows, const uint32_t* ids, size_t n, int8_t* dst,
                        uint32_t* dst_s0, uint32_t* dst_s1, int K, hipStream_t s);
// Synthetic uniformly random shares (bench data; rows rounded up to 16).
void launch_fill_random(int8_t* dst, size_t row0, size_t n_rows, int K, const ChaChaKey& key,
                        uint64_t nonce, uint32_t* sum0, uint32_t* sum1, hipStream_t s);
// Untiles rows back to u16 (tests / fetch).
void launch_untile_rows(const int8_t* src, size_t row0, size_t n_rows, int K, uint16_t* dst,
                        hipStream_t s);

// Block tile (query rows x DB rows) of each compiled GEMM variant.
struct GemmTile {
  int bq, bd;
};
constexpr int kGemmVariants = 2;
// Variants from kRotVariantBase on need a query set with rotation structure
// (uploaded from raw shares, rotations generated on the device).
constexpr int kRotVariantBase = 100, kRotVariants = 3;
// The rotation-aware variant for small query sets (two-query tile).
constexpr int kRotSmallVariant = kRotVariantBase + 2, kRotSmallRows = 2 * 31;
// ... and for sets of up to this many queries whose count is 1 or 2 mod 4
constexpr int kRotSmallAlsoMax = 10;
inline bool is_rot_variant(int v) { return v >= kRotVariantBase && v < kRotVariantBase + kRotVariants; }
constexpr int kMaxBQ = 256;   // query buffers are padded to this many rows
constexpr int kMaxBD = 256;   // DB buffers are padded to this many rows
GemmTile gemm_tile(int variant);
void launch_share_gemm(const GemmArgs& a, int variant, hipStream_t s);
// Persistent form of the rotation-aware 4-query kernel (variant
// kRotVariantBase) over the tiles of n_desc launches: descs and tile_end
// (running tile counts) in device memory, `grid` blocks looping until every
// tile is done. No allocation or synchronisation: capturable.
void launch_share_gemm_multi(const GemmArgs* descs, const uint32_t* tile_end, int n_desc, int grid,
                             hipStream_t s);
// Up to kMaxHeads plain launches of the rotation-aware 4-query kernel
// (variant kRotVariantBase; each head as launch_share_gemm would take it,
// n_seg = 0) as one grid: head i's blocks follow head i - 1's, so the
// dispatcher starts a launch's blocks as the previous launch drains and the
// launches share one kernel boundary. The heads go by value (no device copy):
// capturable, nothing allocated.
constexpr int kMaxHeads = 6;
void launch_share_gemm_heads(const GemmHead* heads, int n, hipStream_t s);
// Padded unrotated limb planes for the rotation-aware variants, built from the
// rotated u16 query rows [qr][K].
size_t rot_planes_bytes(int qr, int K);
void launch_rot_planes(const uint16_t* rows, int qr, int K, int8_t* out, hipStream_t s);

}  // namespace irismpc
