// DistanceComparator kernels: eye merge of match bitmaps, candidate lists of
// the two-stage flow, the index-mapped open, and buffer fills.
//
// mergeDbResults / mergeDbResultsWithOrPolicyBitmap (dot/kernel.cu:241-349):
// AND (or per-pair OR policy) of the left/right bitmaps, atomicMin of the
// first match per query, up to 256 match ids per query, plus the per-eye
// partial lists. Here one thread owns one 64-bit bitmap word and visits only
// its set bits.
#include "comparator.hpp"

#include <mutex>
#include <string>

namespace irismpc {


__global__ __launch_bounds__(256) void merge_kernel(const MergeArgs a) {
  const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = a.query_len * a.db_len;
  if (w * 64 >= total) return;
  const uint64_t L = a.left[w], R = a.right[w];
  uint64_t any = L | R;
  const size_t valid = total - w * 64;
  if (valid < 64) any &= (1ull << valid) - 1;
  while (any) {
    const int i = __builtin_ctzll(any);
    any &= any - 1;
    const size_t bit = w * 64 + (size_t)i;
    const uint32_t q = (uint32_t)(bit / a.db_len), d = (uint32_t)(bit % a.db_len);
    const bool l = (L >> i) & 1, r = (R >> i) & 1;
    if (l) {
      const uint32_t c = atomicAdd(&a.mcl[q], 1u);
      if (c < kAllMatchesLen) a.prl[kAllMatchesLen * q + c] = d;
    }
    if (r) {
      const uint32_t c = atomicAdd(&a.mcr[q], 1u);
      if (c < kAllMatchesLen) a.prr[kAllMatchesLen * q + c] = d;
    }
    bool m = l && r;
    if (a.or_policy) {
      const size_t stride = (a.max_db_len + 63) / 64;
      const bool use_or = (a.or_policy[q * stride + d / 64] >> (d % 64)) & 1;
      m = use_or ? (l || r) : (l && r);
    }
    if (m) {
      atomicMin(&a.final_results[q], d);
      const uint32_t c = atomicAdd(&a.match_counters[q], 1u);
      if (c < kAllMatchesLen) a.all_matches[kAllMatchesLen * q + c] = d;
    }
  }
}

// mergeBatchResults (dot/kernel.cu:352-386) on the compact batch bitmaps:
// one thread per word, visiting the bits set in both eyes.
__global__ __launch_bounds__(256) void merge_batch_kernel(const BatchMergeArgs a) {
  const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = a.batch * a.row_bits;
  if (w * 64 >= total) return;
  uint64_t v = a.left[w] & a.right[w];
  const size_t valid = total - w * 64;
  if (valid < 64) v &= (1ull << valid) - 1;
  while (v) {
    const int i = __builtin_ctzll(v);
    v &= v - 1;
    const size_t bit = w * 64 + (size_t)i;
    const uint32_t q = (uint32_t)(bit / a.row_bits), j = (uint32_t)(bit % a.row_bits);
    if (j >= q) continue;  // only earlier queries of the batch
    atomicMin(&a.final_results[q], 0xFFFFFFFEu);
    const uint32_t c = atomicAdd(&a.match_counters[q], 1u);
    if (c < kAllMatchesLen) a.all_matches[kAllMatchesLen * q + c] = 0xFFFFFFFFu - j;
  }
}

// Exclusive prefix sum over a 256-thread block; *total = the block's sum.
// lds holds the four wave totals (the caller syncs before reusing it).
__device__ inline uint32_t block_exclusive_scan256(uint32_t v, uint32_t* lds, uint32_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wv] = x;
  __syncthreads();
  uint32_t base = 0;
  for (int i = 0; i < wv; i++) base += lds[i];
  *total = lds[0] + lds[1] + lds[2] + lds[3];
  return base + x - v;
}

constexpr int kCandTileWords = 256;  // one 256-thread block per 16,384 ids

// Word w of the candidate set: OR over the batch rows of bitmap bits
// [row_bits * q + 64 w, +64), clipped to db_size.
__global__ __launch_bounds__(256) void cand_union_kernel(const CandidateArgs a) {
  const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w * 64 >= a.db_size) return;
  uint64_t acc = 0;
  for (size_t q = 0; q < a.batch; q++) {
    const size_t o = q * a.row_bits + w * 64;
    const size_t wi = o >> 6;
    const unsigned sh = (unsigned)(o & 63);
    uint64_t v = a.bitmap[wi] >> sh;
    if (sh) v |= a.bitmap[wi + 1] << (64 - sh);
    acc |= v;
  }
  const size_t valid = a.db_size - w * 64;
  if (valid < 64) acc &= (1ull << valid) - 1;
  a.cand[w] = acc;
}

__global__ __launch_bounds__(256) void cand_extra_kernel(const CandidateArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_extra) return;
  const uint32_t d = a.extra[i];
  if (d < a.db_size)
    atomicOr(reinterpret_cast<unsigned long long*>(&a.cand[d >> 6]), 1ull << (d & 63));
}

__global__ __launch_bounds__(256) void cand_count_kernel(const CandidateArgs a) {
  __shared__ uint32_t lds[4];
  const size_t w = (size_t)blockIdx.x * kCandTileWords + threadIdx.x;
  const uint32_t pc = w * 64 < a.db_size ? (uint32_t)__popcll(a.cand[w]) : 0u;
  uint32_t total;
  block_exclusive_scan256(pc, lds, &total);
  if (threadIdx.x == 0) a.tiles[blockIdx.x] = total;
}

// One block: tile counts -> tile start positions, total -> *count.
__global__ __launch_bounds__(256) void cand_scan_kernel(const CandidateArgs a, size_t n_tiles) {
  __shared__ uint32_t lds[4];
  uint32_t carry = 0;
  for (size_t base = 0; base < n_tiles; base += 256) {
    const size_t i = base + threadIdx.x;
    const uint32_t v = i < n_tiles ? a.tiles[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan256(v, lds, &total);
    if (i < n_tiles) a.tiles[i] = carry + ex;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *a.count = carry;
}

__global__ __launch_bounds__(256) void cand_emit_kernel(const CandidateArgs a) {
  __shared__ uint32_t lds[4];
  const size_t w = (size_t)blockIdx.x * kCandTileWords + threadIdx.x;
  uint64_t v = w * 64 < a.db_size ? a.cand[w] : 0ull;
  uint32_t total;
  uint32_t pos = a.tiles[blockIdx.x] + block_exclusive_scan256((uint32_t)__popcll(v), lds, &total);
  while (v) {
    const int i = __builtin_ctzll(v);
    v &= v - 1;
    a.ids[pos++] = (uint32_t)(w * 64 + (size_t)i);
  }
}

__global__ __launch_bounds__(256) void open_mapped_kernel(const MappedOpenArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_words) return;
  uint64_t v = a.a[i] ^ a.b[i] ^ a.c[i];
  const size_t e0 = i * 64;
  const size_t q = e0 / a.chunk_len, d0 = e0 % a.chunk_len;
  if (q >= a.query_len || d0 >= a.real_chunk_len) return;
  const size_t valid = a.real_chunk_len - d0;
  if (valid < 64) v &= (1ull << valid) - 1;
  if (!v) return;
  const size_t slot = a.slot_of ? a.slot_of[q / kRotations] : q / kRotations;
  const uint32_t qi = (uint32_t)(q / kRotations);
  const int8_t rot = (int8_t)((int)(q % kRotations) - kRotations / 2);
  while (v) {
    const int b = __builtin_ctzll(v);
    v &= v - 1;
    const uint32_t id = a.index_map[d0 + (size_t)b];
    if (id >= a.total_db_len) continue;
    if (a.anon.enabled()) anon_store(a.anon, a.dots, e0 + (size_t)b, id, q);
    if (!a.set_bits) continue;
    const size_t pos = a.total_db_len * slot + id;
    atomicOr(reinterpret_cast<unsigned long long*>(&a.bitmap[pos >> 6]), 1ull << (pos & 63));
    const uint32_t c = atomicAdd(a.counter, 1u);
    if (c < a.capacity) {
      a.cap_query[c] = qi;
      a.cap_db[c] = id;
      a.cap_rot[c] = rot;
    }
  }
}

__global__ __launch_bounds__(256) void anon_targets_kernel(const AnonTargetArgs a) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.cap.n_targets * kRotations) return;
  const size_t max_batch = a.cap.max_query_length / kRotations;
  const uint64_t target = a.cap.targets[t / kRotations];
  const size_t db = (size_t)(target / max_batch), j = (size_t)(target % max_batch);
  if (j >= a.batch) return;
  size_t p = 0;
  if (a.index_map) {
    for (p = 0; p < a.real_chunk_len; p++)
      if (a.index_map[p] == db) break;
    if (p == a.real_chunk_len) return;
  } else {
    if (db < a.offset || db >= a.offset + a.real_chunk_len) return;
    p = db - a.offset;
  }
  const size_t q = j * kRotations + t % kRotations;
  const size_t e = q * a.chunk_len + p;
  const uint64_t w = a.a[e >> 6] ^ a.b[e >> 6] ^ a.c[e >> 6];
  if ((w >> (e & 63)) & 1) return;  // a match: captured by the open itself
  anon_store(a.cap, a.dots, e, db, q);
}

void launch_anon_targets(const AnonTargetArgs& a, hipStream_t s) {
  if (!a.cap.enabled() || a.cap.n_targets == 0) return;
  anon_targets_kernel<<<(unsigned)ceil_div(a.cap.n_targets * kRotations, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_u64_kernel(uint64_t* p, size_t n, uint64_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_merge(const MergeArgs& a, hipStream_t s) {
  const size_t words = ceil_div(a.query_len * a.db_len, 64);
  if (!words) return;
  merge_kernel<<<(unsigned)ceil_div(words, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_merge_batch(const BatchMergeArgs& a, hipStream_t s) {
  const size_t words = ceil_div(a.batch * a.row_bits, 64);
  if (!words) return;
  merge_batch_kernel<<<(unsigned)ceil_div(words, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_candidates(const CandidateArgs& a, hipStream_t s) {
  const size_t words = ceil_div(a.db_size, 64);
  const size_t n_tiles = ceil_div(words, kCandTileWords);
  if (words) {
    cand_union_kernel<<<(unsigned)ceil_div(words, 256), 256, 0, s>>>(a);
    IRIS_HIP_CHECK(hipGetLastError());
  }
  if (words && a.n_extra) {
    cand_extra_kernel<<<(unsigned)ceil_div(a.n_extra, 256), 256, 0, s>>>(a);
    IRIS_HIP_CHECK(hipGetLastError());
  }
  if (n_tiles) {
    cand_count_kernel<<<(unsigned)n_tiles, 256, 0, s>>>(a);
    IRIS_HIP_CHECK(hipGetLastError());
  }
  cand_scan_kernel<<<1, 256, 0, s>>>(a, n_tiles);
  IRIS_HIP_CHECK(hipGetLastError());
  if (n_tiles) {
    cand_emit_kernel<<<(unsigned)n_tiles, 256, 0, s>>>(a);
    IRIS_HIP_CHECK(hipGetLastError());
  }
}

// ------------------------------------------------ candidates without allocation
// Scratch of partial_results_noalloc: the per-word candidate bitmap (one u64
// per 64 DB ids) and the per-tile offsets (one u32 per 16,384 ids) of
// launch_candidates, for DB sizes up to kNoAllocMaxDb. Each GPU holds its
// own copy until the process ends. Every call rewrites the words it reads
// (cand_union_kernel writes every candidate word before anything reads it,
// cand_count_kernel every tile), so no call depends on an earlier one's
// contents; calls that overlap in time on one GPU share the array and race.
constexpr size_t kNoAllocCandBytes = kNoAllocMaxDb / 64 * sizeof(uint64_t);
constexpr size_t kNoAllocTileBytes = kNoAllocMaxDb / (64 * kCandTileWords) * sizeof(uint32_t);
static __device__ __align__(256) uint8_t scratch[kNoAllocCandBytes + kNoAllocTileBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

// Device address of this device's `scratch` (looked up once per device: the
// array's name in host code is only a stand-in, not a device address).
static uint8_t* scratch_ptr() {
  int dev = 0;
  IRIS_HIP_CHECK(hipGetDevice(&dev));
  IRIS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    IRIS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    IRIS_REQUIRE(p != nullptr, "no device address for the candidate scratch");
    g_scratch_ptr[dev] = p;
  }
  return static_cast<uint8_t*>(g_scratch_ptr[dev]);
}

void partial_results_noalloc(const uint64_t* bitmap, size_t db_size, size_t batch, size_t row_bits,
                             const uint32_t* extra, size_t n_extra, uint32_t* ids, uint32_t* count,
                             hipStream_t s) {
  IRIS_REQUIRE(db_size <= kNoAllocMaxDb,
               "db_size beyond the candidate scratch (" + std::to_string(kNoAllocMaxDb) + " rows)");
  if (row_bits == 0) row_bits = db_size;
  IRIS_REQUIRE(row_bits >= db_size, "bitmap rows shorter than the DB");
  IRIS_REQUIRE(ids && count && (bitmap || batch == 0 || db_size == 0) && (extra || n_extra == 0),
               "NULL device pointer");
  uint8_t* base = scratch_ptr();
  CandidateArgs c{};
  c.bitmap = bitmap; c.row_bits = row_bits; c.batch = batch; c.db_size = db_size;
  c.extra = extra; c.n_extra = n_extra;
  c.cand = reinterpret_cast<uint64_t*>(base);
  c.tiles = reinterpret_cast<uint32_t*>(base + kNoAllocCandBytes);
  c.ids = ids; c.count = count;
  launch_candidates(c, s);
}

void launch_open_mapped(const MappedOpenArgs& a, hipStream_t s) {
  if (!a.n_words) return;
  open_mapped_kernel<<<(unsigned)ceil_div(a.n_words, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

// Fills are kernels rather than hipMemsetAsync so callers may capture them in
// HIP graphs.
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (!n) return;
  fill_u32_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(p, n, v);
  IRIS_HIP_CHECK(hipGetLastError());
}

void fill_u64(uint64_t* p, size_t n, uint64_t v, hipStream_t s) {
  if (!n) return;
  fill_u64_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(p, n, v);
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
