// Roofline probes for the share GEMM's two ceilings on one GPU.
//
//   roofline_probe mfma   sustained v_mfma_i32_32x32x32_i8 throughput from
//                         registers only (no memory traffic), once with
//                         near-constant operands and once with uniformly random
//                         bytes (what limb shares are): MFMA power, and so the
//                         sustained clock, depends on operand switching.
//   roofline_probe copy   a plain 16-byte-per-lane copy of 2 GiB, the reference
//                         kernel for scaling FETCH_SIZE / WRITE_SIZE counters.
//   roofline_probe read   the HBM read ceiling for a DB-sized buffer (2M rows of
//                         12,800 coefficients x 2 limbs, 53.7 GB), read once
//                         linearly and once in the share GEMM's order (each
//                         wave streams 2 KiB per K-step from four 16-row groups
//                         400 KiB apart), with no MFMA work.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

#define CHECK(x)                                                               \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return 1;                                                                \
    }                                                                          \
  } while (0)

__device__ __forceinline__ int mix(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return (int)x;
}

// 8 waves per block, each with 2 x 2 accumulator tiles and the three limb
// products of the share GEMM per k-slice: 24 MFMAs per iteration.
template <bool RANDOM>
__global__ __launch_bounds__(512, 1) void mfma_loop(int iters, int* out) {
  const int lane = threadIdx.x & 63;
  v4i A[2][2][2], B[2][2][2];
  for (int i = 0; i < 2; i++)
    for (int s = 0; s < 2; s++)
      for (int l = 0; l < 2; l++) {
        const unsigned k = ((blockIdx.x * 512 + threadIdx.x) * 8 + (i * 4 + s * 2 + l)) * 8;
        if (RANDOM) {
          A[i][s][l] = v4i{mix(k), mix(k + 1), mix(k + 2), mix(k + 3)};
          B[i][s][l] = v4i{mix(k + 4), mix(k + 5), mix(k + 6), mix(k + 7)};
        } else {
          A[i][s][l] = v4i{lane + i, s, l, 1};
          B[i][s][l] = v4i{lane - i, l, s, 3};
        }
      }
  v16i acc0[2][2] = {}, acc1[2][2] = {};
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc0[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][s][0], B[j][s][0], acc0[i][j], 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc1[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][s][0], B[j][s][1], acc1[i][j], 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc1[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][s][1], B[j][s][0], acc1[i][j], 0, 0, 0);
    A[0][0][0][0] ^= it;  // keeps the loop from being folded
  }
  int x = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int v = 0; v < 16; v++) x ^= acc0[i][j][v] + acc1[i][j][v];
  out[blockIdx.x * 512 + threadIdx.x] = x;
}

__global__ __launch_bounds__(256) void copy16(const v4i* __restrict__ src, v4i* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ __launch_bounds__(256) void read_linear(const v4i* __restrict__ p, size_t n, v4i* out) {
  v4i x{};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x ^= p[i];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = x;
}

// Block b reads 16-row groups [16 b, 16 b + 16) of the tiled DB layout (group
// stride n_kt x 2 KiB); the lanes of wave w take groups 4 w + 2 j + (lane / 16
// mod 2), 16 bytes at lane offsets as the GEMM's B-operand loads, two K-steps
// per iteration.
__global__ __launch_bounds__(256) void read_groups(const int8_t* __restrict__ db, int n_kt, v4i* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t gstride = (size_t)n_kt * 2048;
  const int r32 = lane & 31;
  const size_t g = (size_t)blockIdx.x * 16 + wave * 4 + (r32 >> 4);
  const int8_t* base = db + g * gstride + ((r32 & 15) + 16 * (lane >> 5)) * 16;
  v4i x{};
  for (int kt = 0; kt < n_kt; kt += 2) {
    v4i b[2][2][4];
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int o = 0; o < 4; o++)
          b[u][j][o] = *reinterpret_cast<const v4i*>(base + 2 * j * gstride + (size_t)(kt + u) * 2048 + o * 512);
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int o = 0; o < 4; o++) x ^= b[u][j][o];
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = x;
}

template <bool RANDOM>
static int time_mfma(const char* name, int blocks, int iters, int reps, int* out) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  mfma_loop<RANDOM><<<blocks, 512>>>(iters, out);  // warm-up and clock ramp
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(a));
  for (int r = 0; r < reps; r++) mfma_loop<RANDOM><<<blocks, 512>>>(iters, out);
  CHECK(hipEventRecord(b));
  CHECK(hipEventSynchronize(b));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, a, b));
  const double ops = (double)reps * blocks * 8 * iters * 24 * (32.0 * 32 * 32 * 2);
  printf("mfma_i32_32x32x32_i8 %-8s operands: %8.1f ms  %7.1f TOPS  (%.1f%% of 5000 TOPS dense int8)\n", name, ms,
         ops / ms / 1e9, ops / ms / 1e9 / 50.0);
  fflush(stdout);
  return 0;
}

int main(int argc, char** argv) {
  const char* mode = argc > 1 ? argv[1] : "mfma";
  if (!strcmp(mode, "mfma")) {
    int* out;
    const int blocks = 2048;
    CHECK(hipMalloc(&out, (size_t)blocks * 512 * 4));
    if (time_mfma<false>("constant", blocks, 200, 400, out)) return 1;
    if (time_mfma<true>("random", blocks, 200, 400, out)) return 1;
    CHECK(hipFree(out));
    return 0;
  }
  if (!strcmp(mode, "copy")) {
    const size_t bytes = (size_t)2 << 30;
    v4i *src, *dst;
    CHECK(hipMalloc(&src, bytes));
    CHECK(hipMalloc(&dst, bytes));
    CHECK(hipMemset(src, 1, bytes));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    copy16<<<8192, 256>>>(src, dst, bytes / 16);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(a));
    copy16<<<8192, 256>>>(src, dst, bytes / 16);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    printf("copy16: %zu bytes read + %zu bytes written per dispatch, %.3f ms, %.1f GB/s\n", bytes, bytes, ms,
           2.0 * bytes / ms / 1e6);
    CHECK(hipFree(src));
    CHECK(hipFree(dst));
    return 0;
  }
  if (!strcmp(mode, "read")) {
    const int n_kt = 200;                             // K = 12,800
    const size_t groups = ((size_t)1 << 21) / 16;     // 2M rows
    const size_t bytes = groups * n_kt * 2048;
    const unsigned blocks_lin = 2048, blocks_grp = (unsigned)(groups / 16);
    int8_t* db;
    v4i* out;
    CHECK(hipMalloc(&db, bytes));
    CHECK(hipMalloc(&out, (size_t)blocks_grp * 256 * sizeof(v4i)));
    CHECK(hipMemset(db, 1, bytes));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    for (int pattern = 0; pattern < 2; pattern++) {
      auto launch = [&] {
        if (pattern == 0)
          read_linear<<<blocks_lin, 256>>>(reinterpret_cast<const v4i*>(db), bytes / 16, out);
        else
          read_groups<<<blocks_grp, 256>>>(db, n_kt, out);
      };
      launch();
      CHECK(hipGetLastError());
      CHECK(hipDeviceSynchronize());
      CHECK(hipEventRecord(a));
      for (int r = 0; r < 3; r++) launch();
      CHECK(hipEventRecord(b));
      CHECK(hipEventSynchronize(b));
      float ms = 0;
      CHECK(hipEventElapsedTime(&ms, a, b));
      ms /= 3;
      printf("read %-28s %zu bytes per dispatch, %.3f ms, %.1f GB/s\n",
             pattern == 0 ? "linear (2048 x 256 threads):" : "share GEMM group order:", bytes, ms,
             bytes / ms / 1e6);
      fflush(stdout);
    }
    CHECK(hipFree(db));
    CHECK(hipFree(out));
    return 0;
  }
  fprintf(stderr, "usage: roofline_probe mfma|copy|read\n");
  return 2;
}
