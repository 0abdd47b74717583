// DistanceComparator kernels: eye merge of match bitmaps and buffer fills.
//
// mergeDbResults / mergeDbResultsWithOrPolicyBitmap (dot/kernel.cu:241-349):
// AND (or per-pair OR policy) of the left/right bitmaps, atomicMin of the
// first match per query, up to 256 match ids per query, plus the per-eye
// partial lists. Here one thread owns one 64-bit bitmap word and visits only
// its set bits.
#include "comparator.hpp"

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
