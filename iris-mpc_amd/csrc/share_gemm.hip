// Share GEMM kernels for gfx950: limb tiling, row sums and the fused
// int8-MFMA GEMM + correction + zero-share-mask epilogue.
//
// Reference behaviour (iris-mpc-gpu/src/dot/share_db.rs:45-112,800-905 and
// dot/kernel.cu:16-28): u16 shares are split into two biased i8 limbs, three
// int8 GEMMs (limb products (0,0), (0,1)*256, (1,0)*256; the (1,1) product
// vanishes mod 2^16) are accumulated in i32, then a correction kernel undoes
// the -128 bias, multiplies mask dots by 2 and adds a ChaCha12 zero-sharing.
// Here all three limb products and the correction/mask epilogue run in one
// kernel; the DB stays resident in HBM in MFMA fragment order.
#include "share_gemm.hpp"

#include <type_traits>
#include "chacha.hpp"

namespace irismpc {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__host__ __device__ inline size_t frag_offset(size_t row_group, int kt, int limb, int n_kt) {
  return (((row_group * (size_t)n_kt + (size_t)kt) * 2 + (size_t)limb) * kFragBytes);
}

// ---------------------------------------------------------------- tiling
// One thread per (row, 16-coefficient group).
__global__ void tile_rows_kernel(const uint16_t* __restrict__ src, size_t n_rows, int K,
                                 int8_t* __restrict__ dst, size_t row0) {
  const int groups = K / 16;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_rows * (size_t)groups) return;
  const size_t r = t / groups;
  const int m = (int)(t % groups);
  const size_t R = row0 + r;
  const int n_kt = K / kKTile;
  const int kt = m / 4, lane = (int)(R % 16) + 16 * (m % 4);
  const uint16_t* s = src + r * (size_t)K + (size_t)m * 16;
  uint4 w0 = reinterpret_cast<const uint4*>(s)[0];
  uint4 w1 = reinterpret_cast<const uint4*>(s)[1];
  uint16_t v[16];
  memcpy(v, &w0, 16);
  memcpy(v + 8, &w1, 16);
  uint8_t lo[16], hi[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    lo[i] = (uint8_t)((v[i] & 0xff) - 128);
    hi[i] = (uint8_t)((v[i] >> 8) - 128);
  }
  int8_t* d0 = dst + frag_offset(R / 16, kt, 0, n_kt) + lane * 16;
  int8_t* d1 = dst + frag_offset(R / 16, kt, 1, n_kt) + lane * 16;
  memcpy(d0, lo, 16);
  memcpy(d1, hi, 16);
}

__global__ void tile_limbs_kernel(const uint8_t* __restrict__ l0, const uint8_t* __restrict__ l1,
                                  size_t n_rows, int K, int8_t* __restrict__ dst, size_t row0) {
  const int groups = K / 16;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_rows * (size_t)groups) return;
  const size_t r = t / groups;
  const int m = (int)(t % groups);
  const size_t R = row0 + r;
  const int n_kt = K / kKTile;
  const int kt = m / 4, lane = (int)(R % 16) + 16 * (m % 4);
  const size_t so = r * (size_t)K + (size_t)m * 16;
  *reinterpret_cast<uint4*>(dst + frag_offset(R / 16, kt, 0, n_kt) + lane * 16) =
      *reinterpret_cast<const uint4*>(l0 + so);
  *reinterpret_cast<uint4*>(dst + frag_offset(R / 16, kt, 1, n_kt) + lane * 16) =
      *reinterpret_cast<const uint4*>(l1 + so);
}

// Serial id of a stored record: big-endian u32 at its start. Row = id - 1;
// records with id 0 or above max_serial are not loaded.
__device__ inline bool record_row(const uint8_t* r, uint32_t max_serial, size_t* row) {
  const uint32_t id = ((uint32_t)r[0] << 24) | ((uint32_t)r[1] << 16) | ((uint32_t)r[2] << 8) | r[3];
  if (id == 0 || id > max_serial) return false;
  *row = id - 1;
  return true;
}

// Stored-record ingest: thread (record, 16-coefficient group) copies the
// group's low-limb bytes (at off_lo in the record) and high-limb bytes (at
// off_hi) into the fragments of row id - 1. Stored bytes are the biased limbs
// already (value - 128 mod 256). Records are only 2-byte aligned: byte loads.
__global__ void tile_records_kernel(const uint8_t* __restrict__ rec, size_t n, size_t stride,
                                    size_t off_lo, size_t off_hi, int K, uint32_t max_serial,
                                    int8_t* __restrict__ dst) {
  const int groups = K / 16;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * (size_t)groups) return;
  const size_t r = t / groups;
  const int m = (int)(t % groups);
  const uint8_t* base = rec + r * stride;
  size_t R;
  if (!record_row(base, max_serial, &R)) return;
  const int n_kt = K / kKTile;
  const int kt = m / 4, lane = (int)(R % 16) + 16 * (m % 4);
  uint8_t lo[16], hi[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    lo[i] = base[off_lo + (size_t)m * 16 + i];
    hi[i] = base[off_hi + (size_t)m * 16 + i];
  }
  memcpy(dst + frag_offset(R / 16, kt, 0, n_kt) + lane * 16, lo, 16);
  memcpy(dst + frag_offset(R / 16, kt, 1, n_kt) + lane * 16, hi, 16);
}

// Sums of the signed limbs (u32 wrapping) of rows [row0, row0 + n_rows):
// one block of four waves per 16-row group. Lane l of a wave holds row l % 16
// in every fragment; wave w reads the fragments of k-tiles w, w + 4, ... (one
// coalesced 1 KiB load per fragment), then the four lanes of a row and the
// four waves are summed. Whole groups are read (buffers are padded to 16-row
// groups); only the rows in range are written.
__global__ void __launch_bounds__(256) row_sums_kernel(const int8_t* __restrict__ tiled, size_t row0,
                                                       size_t n_rows, int K, uint32_t* __restrict__ s0,
                                                       uint32_t* __restrict__ s1) {
  const size_t g = row0 / 16 + blockIdx.x;
  const int w = (int)threadIdx.x / 64, l = (int)threadIdx.x % 64;
  const int n_kt = K / kKTile;
  int a0 = 0, a1 = 0;
#pragma unroll 4
  for (int kt = w; kt < n_kt; kt += 4) {
    const uint4 v0 = *reinterpret_cast<const uint4*>(tiled + frag_offset(g, kt, 0, n_kt) + l * 16);
    const uint4 v1 = *reinterpret_cast<const uint4*>(tiled + frag_offset(g, kt, 1, n_kt) + l * 16);
    const uint32_t x0[4] = {v0.x, v0.y, v0.z, v0.w}, x1[4] = {v1.x, v1.y, v1.z, v1.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
      for (int b = 0; b < 32; b += 8) {
        a0 += (int)(int8_t)(x0[i] >> b);
        a1 += (int)(int8_t)(x1[i] >> b);
      }
    }
  }
  // lanes l, l + 16, l + 32, l + 48 hold row l % 16
  a0 += __shfl_xor(a0, 16);
  a1 += __shfl_xor(a1, 16);
  a0 += __shfl_xor(a0, 32);
  a1 += __shfl_xor(a1, 32);
  __shared__ int sm0[4][16], sm1[4][16];
  if (l < 16) {
    sm0[w][l] = a0;
    sm1[w][l] = a1;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const size_t R = g * 16 + threadIdx.x;
    if (R >= row0 && R < row0 + n_rows) {
      s0[R] = (uint32_t)(sm0[0][threadIdx.x] + sm0[1][threadIdx.x] + sm0[2][threadIdx.x] + sm0[3][threadIdx.x]);
      s1[R] = (uint32_t)(sm1[0][threadIdx.x] + sm1[1][threadIdx.x] + sm1[2][threadIdx.x] + sm1[3][threadIdx.x]);
    }
  }
}

static void launch_row_sums(const int8_t* tiled, size_t row0, size_t n_rows, int K, uint32_t* s0,
                            uint32_t* s1, hipStream_t s) {
  if (n_rows == 0) return;
  const size_t groups = (row0 + n_rows - 1) / 16 - row0 / 16 + 1;
  row_sums_kernel<<<(unsigned)groups, 256, 0, s>>>(tiled, row0, n_rows, K, s0, s1);
  IRIS_HIP_CHECK(hipGetLastError());
}

// Row sums of the rows written by tile_records_kernel (one thread per record).
__global__ void record_sums_kernel(const uint8_t* __restrict__ rec, size_t n, size_t stride,
                                   uint32_t max_serial, const int8_t* __restrict__ tiled, int K,
                                   uint32_t* __restrict__ s0, uint32_t* __restrict__ s1) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  size_t R;
  if (!record_row(rec + r * stride, max_serial, &R)) return;
  const int n_kt = K / kKTile;
  int acc0 = 0, acc1 = 0;
  for (int kt = 0; kt < n_kt; kt++) {
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int lane = (int)(R % 16) + 16 * g;
      int8_t b0[16], b1[16];
      *reinterpret_cast<uint4*>(b0) =
          *reinterpret_cast<const uint4*>(tiled + frag_offset(R / 16, kt, 0, n_kt) + lane * 16);
      *reinterpret_cast<uint4*>(b1) =
          *reinterpret_cast<const uint4*>(tiled + frag_offset(R / 16, kt, 1, n_kt) + lane * 16);
#pragma unroll
      for (int i = 0; i < 16; i++) {
        acc0 += b0[i];
        acc1 += b1[i];
      }
    }
  }
  s0[R] = (uint32_t)acc0;
  s1[R] = (uint32_t)acc1;
}

__global__ void untile_rows_kernel(const int8_t* __restrict__ src, size_t row0, size_t n_rows,
                                   int K, uint16_t* __restrict__ dst) {
  const int groups = K / 16;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_rows * (size_t)groups) return;
  const size_t r = t / groups;
  const int m = (int)(t % groups);
  const size_t R = row0 + r;
  const int n_kt = K / kKTile;
  const int kt = m / 4, lane = (int)(R % 16) + 16 * (m % 4);
  uint8_t lo[16], hi[16];
  *reinterpret_cast<uint4*>(lo) =
      *reinterpret_cast<const uint4*>(src + frag_offset(R / 16, kt, 0, n_kt) + lane * 16);
  *reinterpret_cast<uint4*>(hi) =
      *reinterpret_cast<const uint4*>(src + frag_offset(R / 16, kt, 1, n_kt) + lane * 16);
  uint16_t* d = dst + r * (size_t)K + (size_t)m * 16;
#pragma unroll
  for (int i = 0; i < 16; i++)
    d[i] = (uint16_t)((uint8_t)(lo[i] + 128) | ((uint16_t)(uint8_t)(hi[i] + 128) << 8));
}

// One thread per (16-coefficient group, limb) of the copied row.
__global__ void copy_row_kernel(const int8_t* __restrict__ src, const uint32_t* src_s0,
                                const uint32_t* src_s1, size_t src_row, int8_t* __restrict__ dst,
                                uint32_t* dst_s0, uint32_t* dst_s1, size_t dst_row, int K) {
  const int groups = K / 16;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) {
    dst_s0[dst_row] = src_s0[src_row];
    dst_s1[dst_row] = src_s1[src_row];
  }
  if (t >= 2 * groups) return;
  const int limb = t / groups, m = t % groups;
  const int n_kt = K / kKTile, kt = m / 4;
  const int ls = (int)(src_row % 16) + 16 * (m % 4), ld = (int)(dst_row % 16) + 16 * (m % 4);
  *reinterpret_cast<uint4*>(dst + frag_offset(dst_row / 16, kt, limb, n_kt) + ld * 16) =
      *reinterpret_cast<const uint4*>(src + frag_offset(src_row / 16, kt, limb, n_kt) + ls * 16);
}

// Thread (j, piece): 16-byte piece `piece` (limb, 16-coefficient group) of
// subset row j; consecutive threads take consecutive rows, so the writes of 16
// rows fill one 256-byte stretch of a fragment.
__global__ __launch_bounds__(256) void gather_rows_kernel(
    const int8_t* __restrict__ src, const uint32_t* __restrict__ src_s0,
    const uint32_t* __restrict__ src_s1, size_t src_rows, const uint32_t* __restrict__ ids,
    size_t n, int8_t* __restrict__ dst, uint32_t* dst_s0, uint32_t* dst_s1, int K) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t r = ids[j];
  if (r >= src_rows) return;
  const int groups = K / 16, piece = (int)blockIdx.y;
  if (piece == 0) {
    dst_s0[j] = src_s0[r];
    dst_s1[j] = src_s1[r];
  }
  const int limb = piece / groups, m = piece % groups;
  const int n_kt = K / kKTile, kt = m / 4;
  const int ls = (int)(r % 16) + 16 * (m % 4), ld = (int)(j % 16) + 16 * (m % 4);
  *reinterpret_cast<uint4*>(dst + frag_offset(j / 16, kt, limb, n_kt) + ld * 16) =
      *reinterpret_cast<const uint4*>(src + frag_offset(r / 16, kt, limb, n_kt) + ls * 16);
}

// Uniformly random limbs for rows [row0, row0 + n_rows): byte b of the tiled
// row-group range comes from ChaCha20(key, nonce, counter = offset / 64).
// Uniform u16 shares have exactly the distribution of real Shamir shares.
__global__ void fill_random_kernel(int8_t* __restrict__ dst, size_t byte0, size_t n_bytes,
                                   ChaChaKey key, uint64_t nonce) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t * 64 >= n_bytes) return;
  uint32_t ks[16];
  chacha_block<20>(key, nonce, (byte0 / 64) + t, ks);
  uint4* d = reinterpret_cast<uint4*>(dst + byte0 + t * 64);
#pragma unroll
  for (int i = 0; i < 4; i++) d[i] = uint4{ks[4 * i], ks[4 * i + 1], ks[4 * i + 2], ks[4 * i + 3]};
}

void launch_gather_rows(const int8_t* src, const uint32_t* src_s0, const uint32_t* src_s1,
                        size_t src_rows, const uint32_t* ids, size_t n, int8_t* dst,
                        uint32_t* dst_s0, uint32_t* dst_s1, int K, hipStream_t s) {
  if (!n) return;
  dim3 grid((unsigned)ceil_div(n, 256), (unsigned)(2 * (K / 16)));
  gather_rows_kernel<<<grid, 256, 0, s>>>(src, src_s0, src_s1, src_rows, ids, n, dst, dst_s0,
                                          dst_s1, K);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_fill_random(int8_t* dst, size_t row0, size_t n_rows, int K, const ChaChaKey& key,
                        uint64_t nonce, uint32_t* sum0, uint32_t* sum1, hipStream_t s) {
  IRIS_REQUIRE(row0 % 16 == 0, "random fill starts at a row group");
  const size_t group_bytes = (size_t)K * 2 * 16;
  const size_t b0 = row0 / 16 * group_bytes;
  const size_t nb = ceil_div(n_rows, 16) * group_bytes;
  fill_random_kernel<<<(unsigned)ceil_div(nb / 64, 256), 256, 0, s>>>(dst, b0, nb, key, nonce);
  IRIS_HIP_CHECK(hipGetLastError());
  launch_row_sums(dst, row0, n_rows, K, sum0, sum1, s);
}

void launch_tile_rows(const uint16_t* src, size_t n_rows, int K, int8_t* dst, size_t row0,
                      uint32_t* sum0, uint32_t* sum1, hipStream_t s) {
  if (n_rows == 0) return;
  size_t n = n_rows * (size_t)(K / 16);
  tile_rows_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(src, n_rows, K, dst, row0);
  IRIS_HIP_CHECK(hipGetLastError());
  launch_row_sums(dst, row0, n_rows, K, sum0, sum1, s);
}

void launch_tile_limbs(const uint8_t* l0, const uint8_t* l1, size_t n_rows, int K, int8_t* dst,
                       size_t row0, uint32_t* sum0, uint32_t* sum1, hipStream_t s) {
  if (n_rows == 0) return;
  size_t n = n_rows * (size_t)(K / 16);
  tile_limbs_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(l0, l1, n_rows, K, dst, row0);
  IRIS_HIP_CHECK(hipGetLastError());
  launch_row_sums(dst, row0, n_rows, K, sum0, sum1, s);
}

void launch_tile_records(const uint8_t* records, size_t n, size_t stride, size_t off_lo,
                         size_t off_hi, int K, uint32_t max_serial, int8_t* dst, uint32_t* sum0,
                         uint32_t* sum1, hipStream_t s) {
  if (n == 0) return;
  const size_t t = n * (size_t)(K / 16);
  tile_records_kernel<<<(unsigned)ceil_div(t, 256), 256, 0, s>>>(records, n, stride, off_lo, off_hi,
                                                                 K, max_serial, dst);
  IRIS_HIP_CHECK(hipGetLastError());
  record_sums_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(records, n, stride, max_serial, dst,
                                                                K, sum0, sum1);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_copy_row(const int8_t* src, const uint32_t* src_s0, const uint32_t* src_s1,
                     size_t src_row, int8_t* dst, uint32_t* dst_s0, uint32_t* dst_s1,
                     size_t dst_row, int K, hipStream_t s) {
  copy_row_kernel<<<(unsigned)ceil_div(2 * (K / 16), 256), 256, 0, s>>>(
      src, src_s0, src_s1, src_row, dst, dst_s0, dst_s1, dst_row, K);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_untile_rows(const int8_t* src, size_t row0, size_t n_rows, int K, uint16_t* dst,
                        hipStream_t s) {
  if (n_rows == 0) return;
  size_t n = n_rows * (size_t)(K / 16);
  untile_rows_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(src, row0, n_rows, K, dst);
  IRIS_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- GEMM
// Two i32 accumulators per MFMA tile: acc0 = A0.B0 and acc1 = A0.B1 + A1.B0
// (only acc1 mod 2^8 matters after the << 8).
// v_mfma_i32_16x16x64_i8 operand layout: lane l supplies row (l % 16) and a
// 16-byte k-slice selected by l / 16 for both operands, so any consistent
// k-assignment contracts correctly; the result tile has row 4*(l/16)+r and
// column l % 16 in register r.

// XCD-aware mapping: consecutive block ids land on different XCDs; each XCD
// gets a contiguous run of tiles with the query tile varying fastest so the
// blocks sharing a DB tile run together and hit that XCD's L2.
template <int BQ, int BD>
__device__ __forceinline__ bool block_tile(const GemmArgs& a, int& q0, size_t& db0) {
  const int T = a.n_qtiles * a.n_dtiles;
  const int T8 = (T + 7) / 8;
  const int t = (int)(blockIdx.x & 7) * T8 + (int)(blockIdx.x >> 3);
  if (t >= T) return false;
  const int dtile = t / a.n_qtiles, qtile = t % a.n_qtiles;
  db0 = a.db_begin + (size_t)dtile * BD;
  q0 = qtile * BQ;
  return true;
}

// One K-step of the three limb products. The acc0 pass comes first and each
// acc1 pass covers all tiles, so dependent accumulations into one register
// set are WQ*WD MFMAs apart.
template <int WQ, int WD>
__device__ __forceinline__ void mma_step(v4i (&acc0)[WQ][WD], v4i (&acc1)[WQ][WD],
                                         const v4i (&A)[WQ][2], const v4i (&B)[WD][2]) {
#pragma unroll
  for (int i = 0; i < WQ; i++)
#pragma unroll
    for (int j = 0; j < WD; j++)
      acc0[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][0], B[j][0], acc0[i][j], 0, 0, 0);
#pragma unroll
  for (int i = 0; i < WQ; i++)
#pragma unroll
    for (int j = 0; j < WD; j++)
      acc1[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][0], B[j][1], acc1[i][j], 0, 0, 0);
#pragma unroll
  for (int i = 0; i < WQ; i++)
#pragma unroll
    for (int j = 0; j < WD; j++)
      acc1[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][1], B[j][0], acc1[i][j], 0, 0, 0);
}

// Bias correction into the u16 LDS tile, then multiplier + ChaCha12
// zero-sharing + 64-byte stores (dot/kernel.cu:16-28).
__device__ __forceinline__ uint16_t corrected(const GemmArgs& a, int q, size_t d, uint32_t acc0,
                                               uint32_t acc1) {
  const uint32_t s0 = a.dsum0[d] + a.qsum0[q];
  const uint32_t s1 = a.dsum1[d] + a.qsum1[q];
  return (uint16_t)(acc0 + (acc1 << 8) + (s0 << 7) + ((s0 + s1) << 15));
}

template <int BQ, int BD, class Args>
__device__ __forceinline__ void epilogue_store(const Args& a, const uint16_t* tile, int q0,
                                               size_t db0, int q_end, int tid, int nthreads);

template <int WQ, int WD, int NWQ = 2, int NWD = 2>
__device__ __forceinline__ void gemm_epilogue(const GemmArgs& a, uint16_t* tile,
                                              const v4i (&acc0)[WQ][WD],
                                              const v4i (&acc1)[WQ][WD], int q0, size_t db0,
                                              int wq, int wd, int lane) {
  constexpr int BQ = NWQ * WQ * 16, BD = NWD * WD * 16;
#pragma unroll
  for (int i = 0; i < WQ; i++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ql = (wq * WQ + i) * 16 + 4 * (lane >> 4) + r;
#pragma unroll
      for (int j = 0; j < WD; j++) {
        const int dl = (wd * WD + j) * 16 + (lane & 15);
        tile[ql * BD + dl] = corrected(a, q0 + ql, db0 + dl, (uint32_t)acc0[i][j][r],
                                       (uint32_t)acc1[i][j][r]);
      }
    }
  }
  epilogue_store<BQ, BD>(a, tile, q0, db0, a.qr, (int)threadIdx.x, (int)blockDim.x);
}

// Multiplier, ChaCha12 zero-sharing and 64-byte stores of the u16 tile; query
// rows at or beyond q_end are not stored.
template <int BQ, int BD, class Args>
__device__ __forceinline__ void epilogue_store(const Args& a, const uint16_t* tile, int q0,
                                               size_t db0, int q_end, int tid, int nthreads) {
  __syncthreads();
  constexpr int SEGS = BD / 32;
  for (int s = tid; s < BQ * SEGS; s += nthreads) {
    const int ql = s / SEGS, seg = s % SEGS;
    const int q = q0 + ql;
    const size_t r = db0 + (size_t)seg * 32;
    if (q >= q_end || r >= a.db_end) continue;
    const size_t rl = r - a.db_begin;
    const size_t c = rl / a.chunk;
    const size_t local = rl - c * a.chunk;
    const size_t rem = (a.db_end - a.db_begin) - c * a.chunk;
    const size_t w = rem < a.chunk ? rem : a.chunk;
    uint16_t* out = a.out;
    size_t stride = a.out_chunk_stride;
    uint64_t ctr0 = a.ctr_first, cpc = a.ctr_per_chunk;
    int qq = q;
    if constexpr (Args::kPacked) {
      if (a.n_seg) {  // last segment starting at or before q
        int lo = 0, hi = a.n_seg - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (a.seg[mid].row_begin <= q) lo = mid;
          else hi = mid - 1;
        }
        out = a.seg[lo].out;
        stride = a.seg[lo].out_chunk_stride;
        ctr0 = a.seg[lo].ctr_first;
        cpc = a.seg[lo].ctr_per_chunk;
        qq = q - a.seg[lo].row_begin;
      }
    }
    const size_t idx = (size_t)qq * w + local;
    uint32_t v[16];
    __builtin_memcpy(v, &tile[ql * BD + seg * 32], 64);
    if (a.masked) {
      uint32_t ko[16], kp[16];
      const uint64_t ctr = ctr0 + c * cpc + idx / 32;
      chacha_block<12>(a.key_own, a.nonce, ctr, ko);
      chacha_block<12>(a.key_prev, a.nonce, ctr, kp);
      const uint16_t m = (uint16_t)a.mult;
#pragma unroll
      for (int k = 0; k < 16; k++)  // two u16 results per word, mod 2^16 each
        v[k] = as_u32(as_u16x2(v[k]) * m + as_u16x2(ko[k]) - as_u16x2(kp[k]));
    } else if (a.mult != 1) {
      const uint16_t m = (uint16_t)a.mult;
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = as_u32(as_u16x2(v[k]) * m);
    }
    uint4* dst = reinterpret_cast<uint4*>(out + c * stride + idx);
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k] = uint4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
  }
}

// LDS-staged form: per K-step the block's query and DB fragments (each 1 KiB,
// one coalesced dwordx4 per lane) are fetched from global memory once into a
// double-buffered LDS stage; waves read their operands with ds_read_b128.
// NWQ x NWD waves; with 8 waves two share each SIMD, so one wave's MFMAs
// cover the other's LDS reads.
template <int WQ, int WD, int NWQ, int NWD>
struct LdsStages {
  static constexpr int NA = NWQ * WQ * 2, NB = NWD * WD * 2, NF = NA + NB;
  union {
    v4i stage[2][NF * 64];
    uint16_t tile[(NWQ * WQ * 16) * (NWD * WD * 16)];
  };
};

template <int WQ, int WD, int NWQ, int NWD>
__global__ __launch_bounds__(64 * NWQ * NWD, 1) void share_gemm_lds_kernel(const GemmArgs a) {
  constexpr int NW = NWQ * NWD;
  constexpr int BQ = NWQ * WQ * 16, BD = NWD * WD * 16;
  using S = LdsStages<WQ, WD, NWQ, NWD>;
  constexpr int NA = S::NA, NF = S::NF;
  static_assert(NF % NW == 0, "fragments split evenly over the waves");
  constexpr int PER_WAVE = NF / NW;
  __shared__ S sm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wq = wave / NWD, wd = wave % NWD;
  int q0;
  size_t db0;
  if (!block_tile<BQ, BD>(a, q0, db0)) return;

  const size_t gstride = (size_t)a.n_kt * 2 * kFragBytes;
  const int8_t* qbase = a.q + (size_t)(q0 / 16) * gstride + lane * 16;
  const int8_t* dbase = a.db + (db0 / 16) * gstride + lane * 16;
  // fragment f of a K-step: f < NA -> query group f/2, limb f%2; else DB group
  // (f-NA)/2, limb (f-NA)%2. Wave w fetches fragments w, w+NW, w+2NW, ...
  auto src = [&](int f, int kt) -> const v4i* {
    const size_t ko = (size_t)kt * 2 * kFragBytes;
    if (f < NA) return reinterpret_cast<const v4i*>(qbase + (f >> 1) * gstride + ko + (f & 1) * kFragBytes);
    const int g = f - NA;
    return reinterpret_cast<const v4i*>(dbase + (g >> 1) * gstride + ko + (g & 1) * kFragBytes);
  };

  v4i acc0[WQ][WD], acc1[WQ][WD];
#pragma unroll
  for (int i = 0; i < WQ; i++)
#pragma unroll
    for (int j = 0; j < WD; j++) {
      acc0[i][j] = v4i{0, 0, 0, 0};
      acc1[i][j] = v4i{0, 0, 0, 0};
    }

  v4i stg[PER_WAVE];
#pragma unroll
  for (int m = 0; m < PER_WAVE; m++) stg[m] = *src(wave + NW * m, 0);
#pragma unroll
  for (int m = 0; m < PER_WAVE; m++) sm.stage[0][(wave + NW * m) * 64 + lane] = stg[m];
  __syncthreads();

  for (int kt = 0; kt < a.n_kt; kt++) {
    const int st = kt & 1;
    const bool more = kt + 1 < a.n_kt;
    if (more) {
#pragma unroll
      for (int m = 0; m < PER_WAVE; m++) stg[m] = *src(wave + NW * m, kt + 1);
    }
    v4i A[WQ][2], B[WD][2];
#pragma unroll
    for (int i = 0; i < WQ; i++)
#pragma unroll
      for (int l = 0; l < 2; l++) A[i][l] = sm.stage[st][((wq * WQ + i) * 2 + l) * 64 + lane];
#pragma unroll
    for (int j = 0; j < WD; j++)
#pragma unroll
      for (int l = 0; l < 2; l++) B[j][l] = sm.stage[st][(NA + (wd * WD + j) * 2 + l) * 64 + lane];
    mma_step<WQ, WD>(acc0, acc1, A, B);
    if (more) {
#pragma unroll
      for (int m = 0; m < PER_WAVE; m++) sm.stage[st ^ 1][(wave + NW * m) * 64 + lane] = stg[m];
    }
    __syncthreads();
  }
  gemm_epilogue<WQ, WD, NWQ, NWD>(a, sm.tile, acc0, acc1, q0, db0, wq, wd, lane);
}


// Direct-to-LDS form for gfx950: each fragment (1 KiB, lane l's 16 bytes at
// l * 16) is one wave-wide global_load_lds_dwordx4, which lands in LDS in the
// same lane order. With no register staging the ring can be STAGES deep, so a
// K-step's loads are in flight for STAGES - 1 steps of MFMA work instead of one.
__device__ __forceinline__ void dma_fragment(const int8_t* lane_src, v4i* lds_frag) {
  __builtin_amdgcn_global_load_lds(
      (__attribute__((address_space(1))) void*)(lane_src),
      (__attribute__((address_space(3))) void*)(lds_frag), 16, 0, 0);
}

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding:
// vmcnt bits [3:0] and [15:14], expcnt [6:4] and lgkmcnt [11:8] left at max).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier without the release/acquire fence of __syncthreads (whose
// vmcnt(0) would drain the whole DMA ring): the explicit vmcnt wait before it
// orders this wave's LDS DMA writes, and the empty asm with a memory clobber
// keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}


// Hybrid form: the query tile, which every wave of the block multiplies, is
// shared through a 3-stage LDS DMA ring; each wave's own DB rows bypass LDS
// and load straight into VGPRs. Waves split the DB dimension only (1 x NW),
// each owning TQ x TD 32x32 tiles (accumulators in AGPRs at one wave per
// SIMD). Per MFMA this moves well under half the LDS bytes of the all-LDS
// forms, whose ds_read + DMA traffic saturates LDS.
//
// Step k opens with one group of MFMAs on A(k), B(k) (their first use retires
// every load of step k - 1: with LDS-DMA and register loads both pending the
// compiler waits for vmcnt(0) there), then the barrier that publishes query
// step k + 1, then the DMA of query step k + 2, the DB loads of step k + 1
// and the ds_reads of A(k + 1), then the rest of step k's MFMAs, which cover
// those loads. Loads are unconditional (indices clamped to the last step; the
// surplus DMA lands in a stage nobody reads again).
template <int TQ, int TD, int NW>
struct HybStages {
  static constexpr int NA = TQ * 4;  // query fragments per K-step (16 rows x 64 k x limb)
  static constexpr int STAGES = 3;
  union {
    v4i stage[STAGES][NA * 64];
    uint16_t tile[(TQ * 32) * (NW * TD * 32)];
  };
};

template <int TQ, int TD, int NW>
__global__ __launch_bounds__(64 * NW, 1) void share_gemm_hyb_kernel(const GemmArgs a) {
  constexpr int BQ = TQ * 32, BD = NW * TD * 32;
  using S = HybStages<TQ, TD, NW>;
  constexpr int NA = S::NA, STAGES = S::STAGES;
  static_assert(NA % NW == 0, "query fragments split evenly over the waves");
  constexpr int PER_WAVE = NA / NW;  // DMA instructions per wave per K-step
  constexpr int NMFMA = 2 * 3 * TQ * TD, G1 = TQ * TD;
  __shared__ S sm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int q0;
  size_t db0;
  if (!block_tile<BQ, BD>(a, q0, db0)) return;

  const size_t gstride = (size_t)a.n_kt * 2 * kFragBytes;
  const int r32 = lane & 31, h = lane >> 5, gsub = r32 >> 4;
  const int lane_s0 = (r32 & 15) + 16 * h, lane_s1 = lane_s0 + 32;
  const int8_t* qbase = a.q + (size_t)(q0 / 16) * gstride + lane * 16;
  // lane's DB row: tile j of this wave, 16-row group 2j + gsub, fragment lane lane_s0
  const int8_t* dbase = a.db + (db0 / 16 + (size_t)wave * TD * 2 + gsub) * gstride + lane_s0 * 16;
  const int n_kt = a.n_kt, last = n_kt - 1;

  auto issue_q = [&](int kt, int st) {
    const size_t ko = (size_t)kt * 2 * kFragBytes;
#pragma unroll
    for (int m = 0; m < PER_WAVE; m++) {
      const int f = wave + NW * m;
      dma_fragment(qbase + (f >> 1) * gstride + ko + (f & 1) * kFragBytes, &sm.stage[st][f * 64]);
    }
  };
  auto load_b = [&](v4i (&B)[TD][2][2], int kt) {
    const size_t ko = (size_t)kt * 2 * kFragBytes;
#pragma unroll
    for (int j = 0; j < TD; j++)
#pragma unroll
      for (int sl = 0; sl < 2; sl++)
#pragma unroll
        for (int l = 0; l < 2; l++)
          B[j][sl][l] = *reinterpret_cast<const v4i*>(dbase + (size_t)(2 * j) * gstride + ko +
                                                      l * kFragBytes + sl * 512);
  };
  auto read_a = [&](v4i (&A)[TQ][2][2], int st, int sl) {
    const v4i* stg = sm.stage[st];
#pragma unroll
    for (int i = 0; i < TQ; i++)
#pragma unroll
      for (int l = 0; l < 2; l++)
        A[i][sl][l] = stg[((2 * i + gsub) * 2 + l) * 64 + (sl ? lane_s1 : lane_s0)];
  };

  v16i acc0[TQ][TD], acc1[TQ][TD];
#pragma unroll
  for (int i = 0; i < TQ; i++)
#pragma unroll
    for (int j = 0; j < TD; j++) {
      acc0[i][j] = v16i{};
      acc1[i][j] = v16i{};
    }
  // MFMAs [LO, HI) of one K-step, in the order slice, product, query tile, DB
  // tile: consecutive MFMAs on one accumulator are TQ * TD apart
  auto mfma = [&](const v4i (&A)[TQ][2][2], const v4i (&B)[TD][2][2], auto lo, auto hi) {
    constexpr int LO = decltype(lo)::value, HI = decltype(hi)::value;
#pragma unroll
    for (int sl = 0; sl < 2; sl++)
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int i = 0; i < TQ; i++)
#pragma unroll
          for (int j = 0; j < TD; j++) {
            const int idx = ((sl * 3 + p) * TQ + i) * TD + j;
            if (idx < LO || idx >= HI) continue;
            if (p == 0)
              acc0[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][sl][0], B[j][sl][0], acc0[i][j], 0, 0, 0);
            else if (p == 1)
              acc1[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][sl][0], B[j][sl][1], acc1[i][j], 0, 0, 0);
            else
              acc1[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][sl][1], B[j][sl][0], acc1[i][j], 0, 0, 0);
          }
  };
  using C0 = std::integral_constant<int, 0>;
  using CG = std::integral_constant<int, G1>;
  using CH = std::integral_constant<int, NMFMA / 2>;
  using CN = std::integral_constant<int, NMFMA>;

  // prologue: query steps 0 and 1 in flight with DB(0); drain; A(0)
  issue_q(0, 0);
  issue_q(last < 1 ? last : 1, 1);
  v4i Ba[TD][2][2], Bb[TD][2][2], Aa[TQ][2][2], Ab[TQ][2][2];
  load_b(Ba, 0);
  wait_vmcnt<0>();
  stage_barrier();
  read_a(Aa, 0, 0);
  read_a(Aa, 0, 1);

  // the ds_reads of A(k + 1) go out in two batches of at most 15 so that the
  // lgkmcnt that guards A(k) stays encodable (it would otherwise be 0)
  auto step = [&](int kt, int st, const v4i (&Ac)[TQ][2][2], v4i (&An)[TQ][2][2],
                  const v4i (&Bc)[TD][2][2], v4i (&Bn)[TD][2][2]) {
    mfma(Ac, Bc, C0{}, CG{});
    __builtin_amdgcn_sched_barrier(0);
    wait_vmcnt<0>();
    stage_barrier();
    int s1 = st + 1, s2 = st + 2;
    if (s1 >= STAGES) s1 -= STAGES;
    if (s2 >= STAGES) s2 -= STAGES;
    issue_q(kt + 2 < last ? kt + 2 : last, s2);
    load_b(Bn, kt + 1 < last ? kt + 1 : last);
    read_a(An, s1, 0);
    __builtin_amdgcn_sched_barrier(0);
    mfma(Ac, Bc, CG{}, CH{});
    __builtin_amdgcn_sched_barrier(0);
    read_a(An, s1, 1);
    __builtin_amdgcn_sched_barrier(0);
    mfma(Ac, Bc, CH{}, CN{});
    __builtin_amdgcn_sched_barrier(0);
  };
  int kt = 0, st = 0;
  for (; kt + 1 < n_kt; kt += 2) {
    step(kt, st, Aa, Ab, Ba, Bb);
    st = st + 1 == STAGES ? 0 : st + 1;
    step(kt + 1, st, Ab, Aa, Bb, Ba);
    st = st + 1 == STAGES ? 0 : st + 1;
  }
  if (kt < n_kt) mfma(Aa, Ba, C0{}, CN{});
  __syncthreads();

#pragma unroll
  for (int i = 0; i < TQ; i++)
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int ql = i * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
#pragma unroll
      for (int j = 0; j < TD; j++) {
        const int dl = (wave * TD + j) * 32 + r32;
        sm.tile[ql * BD + dl] = corrected(a, q0 + ql, db0 + dl, (uint32_t)acc0[i][j][v],
                                          (uint32_t)acc1[i][j][v]);
      }
    }
  epilogue_store<BQ, BD>(a, sm.tile, q0, db0, a.qr, (int)threadIdx.x, (int)blockDim.x);
}


// Rotation-aware form. The 31 query rows of one iris are rotations of one
// preprocessed share: row (b, t) holds, in every 800-coefficient row r of the
// share, U_b[r][(c + 4 (15 - t)) mod 800] at coefficient c. A block's query
// tile is TQ whole queries (31 TQ rows in TQ MFMA row tiles, TQ spare rows),
// and their unrotated shares sit in LDS for the block's whole K loop as padded
// limb planes P_b[limb][r][i] = limb(U_b[r][(i - 60) mod 800]), i < 920, so
// rotation t of coefficients [c, c + 16) is the 16 bytes at i = 60 + c +
// 4 (15 - t) (4-byte aligned). The query then costs 29 KiB per query of global
// reads per block instead of 0.8 MiB, the K loop needs no barriers, and only
// DB rows stream from global memory, straight into VGPRs two K-steps ahead.
// TQ = 4 is the MFMA-bound form for large query sets. TQ = 2 serves one or two
// queries, where the scan is bound by DB reads: half the MFMA work of a 4-query
// tile and a register file small enough for two waves per SIMD, so that more
// DB bytes are in flight (6.0 against 5.5 TB/s for one query; a plain read
// kernel reaches 6.2). Plane slots are allocated in groups of kRotQueries.
constexpr int kRotPad = 60, kRotRow = 800 + 2 * kRotPad, kRotQueries = 4;
// Bytes per query slot: the two limb planes of R 800-coefficient rows plus one
// pad dword. A 32-row MFMA tile spans two queries; the lanes of rows
// rho = 31 b + t read the dword at b * stride / 4 + c - t, which falls in bank
// (c - rho) mod 32 only when stride / 4 = 1 (mod 32) -- without the pad
// (stride / 4 = 0 mod 32) the row of one query and the row of the next with the
// same rotation meet in one bank (measured: 43 % of the LDS cycles of the code
// launch were bank conflicts).
__host__ __device__ constexpr int rot_query_stride(int R) { return 2 * R * kRotRow + 4; }

template <int TQ, int TD, int NW>
struct RotLds {
  static constexpr int kMaxRows = kCodeLen / 800;  // 16 rows of 800 coefficients
  // whole 1 KiB DMA fragments: the plane copy may overrun the last plane
  static constexpr int kPlaneBytes =
      (TQ * rot_query_stride(kMaxRows) + kFragBytes - 1) / kFragBytes * kFragBytes;
  union {
    v4i planes[kPlaneBytes / 16];
    uint16_t tile[32 * TQ * (NW * TD * 32)];
  };
};

// One block tile (query rows q0 .., DB rows db0 ..) of the rotation-aware
// form: plane DMA, K loop, epilogue. The caller orders it after every earlier
// use of `sm` by the block.
template <int TQ, int TD, int NW, bool kWpe1, class Args>
__device__ __forceinline__ void rot_tile(const Args& a, int q0, size_t db0, RotLds<TQ, TD, NW>& sm,
                                         int tid) {
  constexpr int BQ = TQ * kRotations, BD = NW * TD * 32;  // 32 TQ rows >= 31 TQ
  constexpr int G = TQ * TD, HALF = 3 * TQ * TD;  // MFMAs per pass, per k-slice
  using S = RotLds<TQ, TD, NW>;
  constexpr int WPE = kWpe1 ? 1 : 2;
  const int lane = tid & 63, wave = tid >> 6;

  const int n_kt = a.n_kt, last = n_kt - 1;
  const int R = n_kt * kKTile / 800;          // 800-coefficient rows of the share
  const int plane = R * kRotRow;              // bytes per limb plane
  const int qbytes = rot_query_stride(R);     // bytes per query slot
  {
    // 1 KiB per wave-wide LDS DMA, all in flight at once (the plane buffer
    // carries 1 KiB of slack for the last partial chunk)
    const int8_t* src = a.qrot + (size_t)(q0 / kRotations) * qbytes + lane * 16;
    const int n_chunks = (TQ * qbytes + kFragBytes - 1) / kFragBytes;
    for (int c = wave; c < n_chunks; c += NW) dma_fragment(src + (size_t)c * kFragBytes, &sm.planes[c * 64]);
    wait_vmcnt<0>();
  }
  __syncthreads();

  const size_t gstride = (size_t)n_kt * 2 * kFragBytes;
  const int r32 = lane & 31, h = lane >> 5, gsub = r32 >> 4;
  const int lane_s0 = (r32 & 15) + 16 * h;
  const int8_t* dbase = a.db + (db0 / 16 + (size_t)wave * TD * 2 + gsub) * gstride + lane_s0 * 16;
  int abase[TQ];  // byte offset of (query, rotation) row of tile i, limb 0, coefficient 0
#pragma unroll
  for (int i = 0; i < TQ; i++) {
    int rho = 32 * i + r32;
    if (rho >= BQ) rho = BQ - 1;  // spare rows: computed, never stored
    const int b = rho / kRotations, t = rho % kRotations;
    abase[i] = b * qbytes + kRotPad + 4 * (15 - t);
  }
  const uint8_t* lds = reinterpret_cast<const uint8_t*>(sm.planes);

  constexpr int NBL = TD * 4, NAR = TQ * 2 * 2;  // DB loads, ds_read2 per slice
  constexpr bool kEarly = NBL <= HALF - G && NAR == 2 * G;
  constexpr bool kThread = NBL <= HALF - G && NAR <= 2 * (HALF - G);
  // In the early-read schedule limb 0 of every query tile comes first: the
  // first two passes of a slice use only limb 0 of A, the third limb 1.
  auto read_a = [&](v4i (&A)[TQ][2], int kt, int sl) {
    const int k0 = kt * kKTile + 32 * sl + 16 * h;
    const int off = k0 + (kRotRow - 800) * (k0 / 800);
#pragma unroll
    for (int x = 0; x < 2 * TQ; x++) {
      const int i = kEarly ? x % TQ : x / 2, l = kEarly ? x / TQ : x % 2;
      const uint32_t* p =
          reinterpret_cast<const uint32_t*>(__builtin_assume_aligned(lds + abase[i] + l * plane + off, 4));
      A[i][l] = v4i{(int)p[0], (int)p[1], (int)p[2], (int)p[3]};
    }
  };
  auto load_b = [&](v4i (&B)[TD][2][2], int kt) {
    const size_t ko = (size_t)kt * 2 * kFragBytes;
#pragma unroll
    for (int j = 0; j < TD; j++)
#pragma unroll
      for (int sl = 0; sl < 2; sl++)
#pragma unroll
        for (int l = 0; l < 2; l++)
          B[j][sl][l] = *reinterpret_cast<const v4i*>(dbase + (size_t)(2 * j) * gstride + ko +
                                                      l * kFragBytes + sl * 512);
  };

  v16i acc0[TQ][TD], acc1[TQ][TD];
#pragma unroll
  for (int i = 0; i < TQ; i++)
#pragma unroll
    for (int j = 0; j < TD; j++) {
      acc0[i][j] = v16i{};
      acc1[i][j] = v16i{};
    }
  // MFMAs [LO, HI) of one k-slice in the order product, query tile, DB tile:
  // consecutive MFMAs on one accumulator are TQ * TD apart
  auto mfma = [&](const v4i (&A)[TQ][2], const v4i (&B)[TD][2][2], int sl, auto lo, auto hi) {
    constexpr int LO = decltype(lo)::value, HI = decltype(hi)::value;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int i = 0; i < TQ; i++)
#pragma unroll
        for (int j = 0; j < TD; j++) {
          const int idx = (p * TQ + i) * TD + j;
          if (idx < LO || idx >= HI) continue;
          if (p == 0)
            acc0[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][0], B[j][sl][0], acc0[i][j], 0, 0, 0);
          else if (p == 1)
            acc1[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][0], B[j][sl][1], acc1[i][j], 0, 0, 0);
          else
            acc1[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][1], B[j][sl][0], acc1[i][j], 0, 0, 0);
        }
  };
  using C0 = std::integral_constant<int, 0>;
  using CG = std::integral_constant<int, G>;
  using CH = std::integral_constant<int, HALF>;

  v4i A0[TQ][2], A1[TQ][2];
  v4i B0[TD][2][2], B1[TD][2][2], B2[TD][2][2];
  load_b(B0, 0);
  load_b(B1, last < 1 ? last : 1);
  read_a(A0, 0, 0);

  // One K-step on B(kt), two k-slices of three passes (limb products) each.
  // With one wave per SIMD (TQ = 4) the next slice's A reads go out threaded
  // two at a time between the MFMAs of a slice's first pass (their registers
  // were last read by the previous slice), so they have two passes to land;
  // slice 0's other two passes carry the DB loads of step kt + 2, one per
  // MFMA (a batch of loads would leave this SIMD's only wave without an MFMA
  // to issue). Measured against issuing the A reads after the first pass:
  // 2 % less time per launch.
  // The two-query tile (two waves per SIMD) has too few MFMAs in a first pass
  // for its reads: each of its slices opens with the first pass (where the
  // compiler waits for the operands), then issues the next loads threaded
  // between the MFMAs of the other two passes.
  static_assert(kEarly || WPE > 1, "one wave per SIMD needs the early-read schedule");
  auto thread_reads = [&]() {
#pragma unroll
    for (int r = 0; r < G; r++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // DS read
    }
  };
  auto step = [&](int kt, const v4i (&Bc)[TD][2][2], v4i (&Bp)[TD][2][2]) {
    if constexpr (kEarly) {
      read_a(A1, kt, 1);
      mfma(A0, Bc, 0, C0{}, CG{});
      thread_reads();
      __builtin_amdgcn_sched_barrier(0);
      load_b(Bp, kt + 2 < last ? kt + 2 : last);
      mfma(A0, Bc, 0, CG{}, CH{});
#pragma unroll
      for (int r = 0; r < HALF - G; r++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);               // MFMA
        if (r < NBL) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // VMEM read
      }
      __builtin_amdgcn_sched_barrier(0);
      read_a(A0, kt + 1 < last ? kt + 1 : last, 0);
      mfma(A1, Bc, 1, C0{}, CG{});
      thread_reads();
      __builtin_amdgcn_sched_barrier(0);
      mfma(A1, Bc, 1, CG{}, CH{});
      __builtin_amdgcn_sched_barrier(0);
    } else {
      mfma(A0, Bc, 0, C0{}, CG{});
      __builtin_amdgcn_sched_barrier(0);
      load_b(Bp, kt + 2 < last ? kt + 2 : last);
      read_a(A1, kt, 1);
      mfma(A0, Bc, 0, CG{}, CH{});
      if constexpr (kThread) {
#pragma unroll
        for (int r = 0; r < HALF - G; r++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                  // MFMA
          if (r < NBL) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);     // VMEM read
          if (r < NAR / 2) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0); // DS read
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      mfma(A1, Bc, 1, C0{}, CG{});
      __builtin_amdgcn_sched_barrier(0);
      read_a(A0, kt + 1 < last ? kt + 1 : last, 0);
      mfma(A1, Bc, 1, CG{}, CH{});
      if constexpr (kThread) {
#pragma unroll
        for (int r = 0; r < HALF - G; r++) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                  // MFMA
          if (r < NAR / 2) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0); // DS read
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  int kt = 0;
  for (; kt + 2 < n_kt; kt += 3) {
    step(kt, B0, B2);
    step(kt + 1, B1, B0);
    step(kt + 2, B2, B1);
  }
  if (kt < n_kt) step(kt, B0, B2);
  if (kt + 1 < n_kt) step(kt + 1, B1, B0);
  __syncthreads();

  // The limb-sum correction splits into a query term and a DB term:
  // (s0 << 7) + ((s0 + s1) << 15) with s = dsum + qsum is linear in u32.
  // Query terms of the tile's 32 TQ rows go to LDS past the output tile.
  constexpr int ROWS = 32 * TQ;
  const int q_end = a.qr < q0 + BQ ? a.qr : q0 + BQ;
  uint32_t* qterm = reinterpret_cast<uint32_t*>(&sm.tile[ROWS * BD]);
  static_assert(sizeof(S) >= ROWS * BD * 2 + ROWS * 4, "query terms fit behind the tile");
  static_assert(ROWS <= 64 * NW, "one thread per query term");
  if (tid < ROWS) {
    const int q = q0 + tid;
    uint32_t t = 0;
    if (q < q_end) {
      const uint32_t s0 = a.qsum0[q], s1 = a.qsum1[q];
      t = (s0 << 7) + ((s0 + s1) << 15);
    }
    qterm[tid] = t;
  }
  uint32_t dterm[TD];
#pragma unroll
  for (int j = 0; j < TD; j++) {
    const size_t d = db0 + (wave * TD + j) * 32 + r32;
    const uint32_t s0 = a.dsum0[d], s1 = a.dsum1[d];
    dterm[j] = (s0 << 7) + ((s0 + s1) << 15);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < TQ; i++)
#pragma unroll
    for (int v4 = 0; v4 < 4; v4++) {
      const int qb = i * 32 + 8 * v4 + 4 * h;  // rows qb .. qb + 3 in registers 4 v4 ..
      uint32_t qt[4];
      __builtin_memcpy(qt, &qterm[qb], 16);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int v = 4 * v4 + r, ql = qb + r;
        if (q0 + ql >= q_end) continue;
#pragma unroll
        for (int j = 0; j < TD; j++) {
          const int dl = (wave * TD + j) * 32 + r32;
          sm.tile[ql * BD + dl] = (uint16_t)((uint32_t)acc0[i][j][v] + ((uint32_t)acc1[i][j][v] << 8) +
                                             qt[r] + dterm[j]);
        }
      }
    }
  epilogue_store<ROWS, BD>(a, sm.tile, q0, db0, q_end, tid, 64 * NW);
}

template <int TQ, int TD, int NW, int WPE>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void share_gemm_rot_kernel(const GemmArgs a) {
  constexpr int BQ = TQ * kRotations, BD = NW * TD * 32;
  __shared__ RotLds<TQ, TD, NW> sm;
  int q0;
  size_t db0;
  if (!block_tile<BQ, BD>(a, q0, db0)) return;
  rot_tile<TQ, TD, NW, WPE == 1>(a, q0, db0, sm, (int)threadIdx.x);
}

// Persistent form over several launches' worth of tiles: descs[0 .. n_desc)
// are the arguments of consecutive packed launches (the packs of one chunk),
// tile_end[i] the running tile count through desc i. Each block loops over
// the tiles t = x T8 + j, j + G/8, ... of its XCD's contiguous run (x = block
// mod 8, G blocks), so the blocks sharing a DB tile run together on one XCD as
// in the one-launch-per-pack form; every block leaves once past the run.
template <int TQ, int TD, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(1, 1)))
void share_gemm_rot_multi_kernel(const GemmArgs* __restrict__ descs, const uint32_t* __restrict__ tile_end,
                                 int n_desc) {
  constexpr int BQ = TQ * kRotations, BD = NW * TD * 32;
  __shared__ RotLds<TQ, TD, NW> sm;
  const uint32_t T = tile_end[n_desc - 1];
  const uint32_t T8 = (T + 7) / 8, x = blockIdx.x & 7, per_xcd = gridDim.x >> 3;
  const uint32_t run_end = (x + 1) * T8 < T ? (x + 1) * T8 : T;
  for (uint32_t t = x * T8 + (blockIdx.x >> 3); t < run_end; t += per_xcd) {
    int d = 0;
    while (d + 1 < n_desc && tile_end[d] <= t) d++;  // at most 64 descriptors
    d = __builtin_amdgcn_readfirstlane(d);           // block-uniform: scalar loads of the descriptor
    const GemmArgs& a = descs[d];
    const uint32_t local = (uint32_t)__builtin_amdgcn_readfirstlane((int)(t - (d ? tile_end[d - 1] : 0u)));
    const int q0 = (int)(local % (uint32_t)a.n_qtiles) * BQ;
    const size_t db0 = a.db_begin + (size_t)(local / (uint32_t)a.n_qtiles) * BD;
    __syncthreads();  // the previous tile's epilogue is done with the LDS
    // the thread index, opaque per tile: keeps the compiler from hoisting the
    // tile's per-lane address arithmetic out of the loop (it would stay live
    // through the K loop and spill)
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    rot_tile<TQ, TD, NW, true>(a, q0, db0, sm, tid);
  }
}

// Several plain launches of the 4-query kernel as one grid: blocks
// [block_end[i - 1], block_end[i]) are launch i's, each segment a multiple of
// 8 blocks, so a block's XCD (block mod 8) and its tile are those of the plain
// launch (block_tile on the block's index within its segment).
struct GemmHeads {
  GemmHead h[kMaxHeads];
  uint32_t block_end[kMaxHeads];
  int n;
};

template <int TQ, int TD, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(1, 1)))
void share_gemm_rot_heads_kernel(const GemmHeads hs) {
  constexpr int BQ = TQ * kRotations, BD = NW * TD * 32;
  __shared__ RotLds<TQ, TD, NW> sm;
  const uint32_t b = blockIdx.x;
  int d = 0;
  while (d + 1 < hs.n && hs.block_end[d] <= b) d++;
  const GemmHead& a = hs.h[d];
  const uint32_t local = b - (d ? hs.block_end[d - 1] : 0u);
  const int T = a.n_qtiles * a.n_dtiles, T8 = (T + 7) / 8;
  const int t = (int)(local & 7) * T8 + (int)(local >> 3);
  if (t >= T) return;
  const int dtile = t / a.n_qtiles, qtile = t % a.n_qtiles;
  rot_tile<TQ, TD, NW, true>(a, qtile * BQ, a.db_begin + (size_t)dtile * BD, sm, (int)threadIdx.x);
}

// Padded unrotated limb planes of every query slot from the rotated u16 rows
// (rotation 15 is the unrotated share; slots past the rows are zero shares).
__global__ __launch_bounds__(256) void rot_planes_kernel(const uint16_t* rows, int qr, int K,
                                                         int slots, int8_t* out) {
  const int R = K / 800;
  const size_t n = (size_t)slots * R * kRotRow;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % kRotRow);
  const int r = (int)((t / kRotRow) % R);
  const int b = (int)(t / ((size_t)kRotRow * R));
  const int src_row = b * kRotations + 15;
  uint16_t v = 0;
  if (src_row < qr) v = rows[(size_t)src_row * K + r * 800 + (i - kRotPad + 800) % 800];
  int8_t* q = out + (size_t)b * rot_query_stride(R) + (size_t)r * kRotRow + i;
  q[0] = (int8_t)((v & 0xff) ^ 0x80);
  q[(size_t)R * kRotRow] = (int8_t)((v >> 8) ^ 0x80);
}

size_t rot_planes_bytes(int qr, int K) {
  const size_t slots = ceil_div((size_t)qr, (size_t)(kRotQueries * kRotations)) * kRotQueries;
  return slots * (size_t)rot_query_stride(K / 800) + kFragBytes;
}

void launch_rot_planes(const uint16_t* rows, int qr, int K, int8_t* out, hipStream_t s) {
  IRIS_REQUIRE(K % 800 == 0 && K <= kCodeLen, "rotation planes need K a multiple of 800");
  const int slots = (int)(ceil_div((size_t)qr, (size_t)(kRotQueries * kRotations)) * kRotQueries);
  const size_t n = (size_t)slots * (K / 800) * kRotRow;
  rot_planes_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(rows, qr, K, slots, out);
  IRIS_HIP_CHECK(hipGetLastError());
}

GemmTile gemm_tile(int variant) {
  switch (variant) {
    case 0: return {128, 256};   // 16x16x64 MFMA, LDS-staged, 8 waves (2 per SIMD)
    case 1: return {128, 256};   // 32x32x32 MFMA, query via LDS DMA, DB direct, 4 waves of 128x64
    case kRotVariantBase: return {4 * kRotations, 256};      // rotation-aware, 4 waves
    case kRotVariantBase + 1: return {4 * kRotations, 128};  // rotation-aware, 2 waves
    case kRotVariantBase + 2: return {2 * kRotations, 128};  // two queries, 2 waves per SIMD
    default: throw Error("unknown share GEMM variant " + std::to_string(variant));
  }
}

void launch_share_gemm_multi(const GemmArgs* descs, const uint32_t* tile_end, int n_desc, int grid,
                             hipStream_t s) {
  IRIS_REQUIRE(n_desc >= 1 && n_desc <= kMaxSegments, "1 .. 64 launches per persistent launch");
  IRIS_REQUIRE(grid >= 8 && grid % 8 == 0, "persistent grid: a multiple of 8 blocks");
  share_gemm_rot_multi_kernel<4, 2, 4><<<(unsigned)grid, 256, 0, s>>>(descs, tile_end, n_desc);
  IRIS_HIP_CHECK(hipGetLastError());
}

// The geometry checks of a plain launch.
static void check_gemm_args(const GemmHead& a, int variant, int n_seg) {
  const GemmTile t = gemm_tile(variant);
  IRIS_REQUIRE(a.db_begin % 16 == 0, "db_begin must be a multiple of 16");
  IRIS_REQUIRE((a.db_end - a.db_begin) % 64 == 0 && a.db_end > a.db_begin, "bad DB range");
  IRIS_REQUIRE(a.chunk % 256 == 0, "chunk must be a multiple of 256 rows");
  IRIS_REQUIRE(a.n_qtiles == (int)ceil_div((size_t)a.qr, (size_t)t.bq), "n_qtiles mismatch");
  IRIS_REQUIRE(!is_rot_variant(variant) || (a.qrot != nullptr && (a.n_kt * kKTile) % 800 == 0),
               "rotation-aware variant needs a query uploaded from raw shares");
  IRIS_REQUIRE(n_seg >= 0 && n_seg <= kMaxSegments, "too many batches in one packed launch");
  IRIS_REQUIRE(a.n_dtiles == (int)ceil_div(a.db_end - a.db_begin, (size_t)t.bd), "n_dtiles mismatch");
}

void launch_share_gemm_heads(const GemmHead* heads, int n, hipStream_t s) {
  IRIS_REQUIRE(n >= 1 && n <= kMaxHeads, "1 .. 6 launches per joint launch");
  GemmHeads hs{};
  uint32_t blocks = 0;
  for (int i = 0; i < n; i++) {
    check_gemm_args(heads[i], kRotVariantBase, 0);
    const uint32_t T = (uint32_t)(heads[i].n_qtiles * heads[i].n_dtiles);
    hs.h[i] = heads[i];
    blocks += 8 * ((T + 7) / 8);
    hs.block_end[i] = blocks;
  }
  hs.n = n;
  share_gemm_rot_heads_kernel<4, 2, 4><<<blocks, 256, 0, s>>>(hs);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_share_gemm(const GemmArgs& a, int variant, hipStream_t s) {
  check_gemm_args(a, variant, a.n_seg);
  const int T = a.n_qtiles * a.n_dtiles;
  const unsigned grid = (unsigned)(8 * ((T + 7) / 8));
  switch (variant) {
    case 0: share_gemm_lds_kernel<4, 4, 2, 4><<<grid, 512, 0, s>>>(a); break;
    case 1: share_gemm_hyb_kernel<4, 2, 4><<<grid, 256, 0, s>>>(a); break;
    case kRotVariantBase: share_gemm_rot_kernel<4, 2, 4, 1><<<grid, 256, 0, s>>>(a); break;
    case kRotVariantBase + 1: share_gemm_rot_kernel<4, 2, 2, 1><<<grid, 128, 0, s>>>(a); break;
    case kRotVariantBase + 2: share_gemm_rot_kernel<2, 1, 4, 2><<<grid, 256, 0, s>>>(a); break;
  }
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
