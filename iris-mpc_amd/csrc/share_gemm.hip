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
#include "chacha.hpp"

namespace irismpc {

typedef int v4i __attribute__((ext_vector_type(4)));

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

// One thread per row: sums of the signed limbs (u32 wrapping).
__global__ void row_sums_kernel(const int8_t* __restrict__ tiled, size_t row0, size_t n_rows,
                                int K, uint32_t* __restrict__ s0, uint32_t* __restrict__ s1) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const size_t R = row0 + r;
  const int n_kt = K / kKTile;
  int acc0 = 0, acc1 = 0;
  for (int kt = 0; kt < n_kt; kt++) {
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int lane = (int)(R % 16) + 16 * g;
      const int8_t* p0 = tiled + frag_offset(R / 16, kt, 0, n_kt) + lane * 16;
      const int8_t* p1 = tiled + frag_offset(R / 16, kt, 1, n_kt) + lane * 16;
      int8_t b0[16], b1[16];
      *reinterpret_cast<uint4*>(b0) = *reinterpret_cast<const uint4*>(p0);
      *reinterpret_cast<uint4*>(b1) = *reinterpret_cast<const uint4*>(p1);
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

void launch_fill_random(int8_t* dst, size_t row0, size_t n_rows, int K, const ChaChaKey& key,
                        uint64_t nonce, uint32_t* sum0, uint32_t* sum1, hipStream_t s) {
  IRIS_REQUIRE(row0 % 16 == 0, "random fill starts at a row group");
  const size_t group_bytes = (size_t)K * 2 * 16;
  const size_t b0 = row0 / 16 * group_bytes;
  const size_t nb = ceil_div(n_rows, 16) * group_bytes;
  fill_random_kernel<<<(unsigned)ceil_div(nb / 64, 256), 256, 0, s>>>(dst, b0, nb, key, nonce);
  IRIS_HIP_CHECK(hipGetLastError());
  row_sums_kernel<<<(unsigned)ceil_div(n_rows, 256), 256, 0, s>>>(dst, row0, n_rows, K, sum0, sum1);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_tile_rows(const uint16_t* src, size_t n_rows, int K, int8_t* dst, size_t row0,
                      uint32_t* sum0, uint32_t* sum1, hipStream_t s) {
  if (n_rows == 0) return;
  size_t n = n_rows * (size_t)(K / 16);
  tile_rows_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(src, n_rows, K, dst, row0);
  IRIS_HIP_CHECK(hipGetLastError());
  row_sums_kernel<<<(unsigned)ceil_div(n_rows, 256), 256, 0, s>>>(dst, row0, n_rows, K, sum0, sum1);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_tile_limbs(const uint8_t* l0, const uint8_t* l1, size_t n_rows, int K, int8_t* dst,
                       size_t row0, uint32_t* sum0, uint32_t* sum1, hipStream_t s) {
  if (n_rows == 0) return;
  size_t n = n_rows * (size_t)(K / 16);
  tile_limbs_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(l0, l1, n_rows, K, dst, row0);
  IRIS_HIP_CHECK(hipGetLastError());
  row_sums_kernel<<<(unsigned)ceil_div(n_rows, 256), 256, 0, s>>>(dst, row0, n_rows, K, sum0, sum1);
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
// Block = 4 waves (2 along queries x 2 along DB rows). Wave tile = WQ x WD
// MFMA 16x16 tiles; two i32 accumulators per tile: acc0 = A0.B0 and
// acc1 = A0.B1 + A1.B0 (only acc1 mod 2^8 matters after the << 8).
// v_mfma_i32_16x16x64_i8 operand layout: lane l supplies row (l % 16) and a
// 16-byte k-slice selected by l / 16 for both operands, so any consistent
// k-assignment contracts correctly; the result tile has row 4*(l/16)+r and
// column l % 16 in register r.
template <int WQ, int WD>
__global__ __launch_bounds__(256, 1) void share_gemm_kernel(const GemmArgs a) {
  constexpr int BQ = 2 * WQ * 16, BD = 2 * WD * 16;
  __shared__ uint16_t tile[BQ * BD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wq = wave >> 1, wd = wave & 1;

  // XCD-aware mapping: consecutive block ids land on different XCDs; give
  // each XCD a contiguous run of tiles with the query tile varying fastest so
  // the blocks sharing a DB tile run together and hit that XCD's L2.
  const int T = a.n_qtiles * a.n_dtiles;
  const int T8 = (T + 7) / 8;
  const int t = (int)(blockIdx.x & 7) * T8 + (int)(blockIdx.x >> 3);
  if (t >= T) return;
  const int dtile = t / a.n_qtiles, qtile = t % a.n_qtiles;
  const size_t db0 = a.db_begin + (size_t)dtile * BD;
  const int q0 = qtile * BQ;

  const size_t gstride = (size_t)a.n_kt * 2 * kFragBytes;  // bytes per 16-row group
  const int8_t* qp = a.q + (size_t)(q0 / 16 + wq * WQ) * gstride + lane * 16;
  const int8_t* dp = a.db + (db0 / 16 + (size_t)wd * WD) * gstride + lane * 16;

  v4i acc0[WQ][WD], acc1[WQ][WD];
#pragma unroll
  for (int i = 0; i < WQ; i++)
#pragma unroll
    for (int j = 0; j < WD; j++) {
      acc0[i][j] = v4i{0, 0, 0, 0};
      acc1[i][j] = v4i{0, 0, 0, 0};
    }

  v4i A0[WQ][2], B0[WD][2], A1[WQ][2], B1[WD][2];
  auto load = [&](v4i (&A)[WQ][2], v4i (&B)[WD][2], int kt) {
    const size_t ko = (size_t)kt * 2 * kFragBytes;
#pragma unroll
    for (int i = 0; i < WQ; i++) {
      A[i][0] = *reinterpret_cast<const v4i*>(qp + i * gstride + ko);
      A[i][1] = *reinterpret_cast<const v4i*>(qp + i * gstride + ko + kFragBytes);
    }
#pragma unroll
    for (int j = 0; j < WD; j++) {
      B[j][0] = *reinterpret_cast<const v4i*>(dp + j * gstride + ko);
      B[j][1] = *reinterpret_cast<const v4i*>(dp + j * gstride + ko + kFragBytes);
    }
  };
  auto mma = [&](const v4i (&A)[WQ][2], const v4i (&B)[WD][2]) {
#pragma unroll
    for (int i = 0; i < WQ; i++)
#pragma unroll
      for (int j = 0; j < WD; j++) {
        acc0[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][0], B[j][0], acc0[i][j], 0, 0, 0);
        acc1[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][0], B[j][1], acc1[i][j], 0, 0, 0);
        acc1[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[i][1], B[j][0], acc1[i][j], 0, 0, 0);
      }
  };

  load(A0, B0, 0);
  int kt = 0;
  for (; kt + 2 <= a.n_kt; kt += 2) {
    load(A1, B1, kt + 1);
    mma(A0, B0);
    if (kt + 2 < a.n_kt) load(A0, B0, kt + 2);
    mma(A1, B1);
  }
  if (kt < a.n_kt) mma(A0, B0);

  // ---- epilogue 1: bias correction -> u16 tile in LDS
#pragma unroll
  for (int i = 0; i < WQ; i++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ql = (wq * WQ + i) * 16 + 4 * (lane >> 4) + r;
      const uint32_t qs0 = a.qsum0[q0 + ql], qs1 = a.qsum1[q0 + ql];
#pragma unroll
      for (int j = 0; j < WD; j++) {
        const int dl = (wd * WD + j) * 16 + (lane & 15);
        const uint32_t s0 = a.dsum0[db0 + dl] + qs0;
        const uint32_t s1 = a.dsum1[db0 + dl] + qs1;
        const uint32_t c = (uint32_t)acc0[i][j][r] + ((uint32_t)acc1[i][j][r] << 8);
        tile[ql * BD + dl] = (uint16_t)(c + (s0 << 7) + ((s0 + s1) << 15));
      }
    }
  }
  __syncthreads();

  // ---- epilogue 2: multiplier, ChaCha12 zero-sharing, 64-byte stores
  constexpr int SEGS = BD / 32;
  for (int s = threadIdx.x; s < BQ * SEGS; s += 256) {
    const int ql = s / SEGS, seg = s % SEGS;
    const int q = q0 + ql;
    const size_t r = db0 + (size_t)seg * 32;
    if (q >= a.qr || r >= a.db_end) continue;
    const size_t rl = r - a.db_begin;
    const size_t c = rl / a.chunk;
    const size_t local = rl - c * a.chunk;
    const size_t rem = (a.db_end - a.db_begin) - c * a.chunk;
    const size_t w = rem < a.chunk ? rem : a.chunk;
    const size_t idx = (size_t)q * w + local;
    uint32_t v[16];
    __builtin_memcpy(v, &tile[ql * BD + seg * 32], 64);
    if (a.masked) {
      uint32_t ko[16], kp[16];
      const uint64_t ctr = a.ctr_first + c * a.ctr_per_chunk + idx / 32;
      chacha_block<12>(a.key_own, a.nonce, ctr, ko);
      chacha_block<12>(a.key_prev, a.nonce, ctr, kp);
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const uint32_t lo = (v[k] & 0xffff) * a.mult + (ko[k] & 0xffff) - (kp[k] & 0xffff);
        const uint32_t hi = (v[k] >> 16) * a.mult + (ko[k] >> 16) - (kp[k] >> 16);
        v[k] = (lo & 0xffff) | (hi << 16);
      }
    } else if (a.mult != 1) {
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const uint32_t lo = (v[k] & 0xffff) * a.mult, hi = (v[k] >> 16) * a.mult;
        v[k] = (lo & 0xffff) | (hi << 16);
      }
    }
    uint4* dst = reinterpret_cast<uint4*>(a.out + c * a.out_chunk_stride + idx);
#pragma unroll
    for (int k = 0; k < 4; k++) dst[k] = uint4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
    if (a.out_peer) {
      uint4* pd = reinterpret_cast<uint4*>(a.out_peer + c * a.out_chunk_stride + idx);
#pragma unroll
      for (int k = 0; k < 4; k++) pd[k] = uint4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
    }
  }
}

GemmTile gemm_tile(int variant) {
  switch (variant) {
    case 0: return {128, 128};
    case 1: return {64, 256};
    case 2: return {128, 256};
    case 3: return {64, 128};
    default: throw Error("unknown share GEMM variant " + std::to_string(variant));
  }
}

void launch_share_gemm(const GemmArgs& a, int variant, hipStream_t s) {
  const GemmTile t = gemm_tile(variant);
  IRIS_REQUIRE(a.db_begin % 16 == 0, "db_begin must be a multiple of 16");
  IRIS_REQUIRE((a.db_end - a.db_begin) % 64 == 0 && a.db_end > a.db_begin, "bad DB range");
  IRIS_REQUIRE(a.chunk % 256 == 0, "chunk must be a multiple of 256 rows");
  IRIS_REQUIRE(a.n_qtiles == (int)ceil_div((size_t)a.qr, (size_t)t.bq), "n_qtiles mismatch");
  IRIS_REQUIRE(a.n_dtiles == (int)ceil_div(a.db_end - a.db_begin, (size_t)t.bd), "n_dtiles mismatch");
  const int T = a.n_qtiles * a.n_dtiles;
  const unsigned grid = (unsigned)(8 * ((T + 7) / 8));
  switch (variant) {
    case 0: share_gemm_kernel<4, 4><<<grid, 256, 0, s>>>(a); break;
    case 1: share_gemm_kernel<2, 8><<<grid, 256, 0, s>>>(a); break;
    case 2: share_gemm_kernel<4, 8><<<grid, 256, 0, s>>>(a); break;
    case 3: share_gemm_kernel<2, 4><<<grid, 256, 0, s>>>(a); break;
  }
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
