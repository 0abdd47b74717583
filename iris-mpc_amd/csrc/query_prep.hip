// Query preprocessing on the GPU: Lagrange interpolation factor, basis B and
// the 31 rotations of every query share (the host-side rayon work of the
// reference: galois_engine.rs:23-52,479-488 via server/mod.rs:196-285).
//
// Rotation t of a share = rotate_coefs_left(16) then (t+1) x rotate_right(1)
// inside every 800-coefficient row, i.e. a left rotation by (15 - t) GR
// elements (4 coefficients each) of every 200-element row. Preprocessing is
// per GR element, so it commutes with the rotation.
//
// Mirrored queries (the Mirror orientation, galois_engine.rs:218-224,276-289,
// 518-545): element column c of every row is taken from column 99 - c
// (c < 100) or 299 - c (c >= 100) of the raw share before rotating; for
// codes the imaginary half (rows 8..15, coefficients 6400..) is negated. The
// reference negates by subtracting from a fixed sharing of zero; the plain
// negation shares the same secret, so the opened results are the same.
#include "galois.hpp"
#include "query_prep.hpp"

namespace irismpc {

struct QueryPrepArgs {
  const uint16_t* src;  // batch x K raw query shares of one party
  uint16_t* dst;        // rows_out x K preprocessed rotated rows
  int K;
  int batch;
  int rows_out;
  int mirror;  // 0: as given; 1: mirrored (masks); 2: mirrored, imaginary half negated (codes)
  uint16_t lag[4];
  uint16_t to_b[16];
};

__global__ __launch_bounds__(256) void query_prep_kernel(const QueryPrepArgs a) {
  const int per_row = a.K / 4;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)a.rows_out * per_row) return;
  const int r = (int)(t / per_row), g = (int)(t % per_row);
  const int b = r / kRotations, rot = r % kRotations;
  uint2* out = reinterpret_cast<uint2*>(a.dst + (size_t)r * a.K + 4 * (size_t)g);
  if (b >= a.batch) {
    *out = uint2{0, 0};
    return;
  }
  const int row = g / 200, e = g % 200;
  int src_e = (e + 15 - rot + 200) % 200;
  if (a.mirror) src_e = src_e < 100 ? 99 - src_e : 299 - src_e;
  const uint2 raw = *reinterpret_cast<const uint2*>(a.src + (size_t)b * a.K + 4 * (size_t)(row * 200 + src_e));
  uint32_t x0 = raw.x & 0xffff, x1 = raw.x >> 16, x2 = raw.y & 0xffff, x3 = raw.y >> 16;
  if (a.mirror == 2 && 4 * g >= a.K / 2) {  // products are taken mod 2^16 below
    x0 = 0x10000u - x0;
    x1 = 0x10000u - x1;
    x2 = 0x10000u - x2;
    x3 = 0x10000u - x3;
  }
  const uint32_t l0 = a.lag[0], l1 = a.lag[1], l2 = a.lag[2], l3 = a.lag[3];
  // monomial product with x^4 = x + 1
  const uint32_t d4 = x1 * l3 + x2 * l2 + x3 * l1, d5 = x2 * l3 + x3 * l2, d6 = x3 * l3;
  const uint32_t p[4] = {x0 * l0 + d4, x0 * l1 + x1 * l0 + d4 + d5,
                         x0 * l2 + x1 * l1 + x2 * l0 + d5 + d6,
                         x0 * l3 + x1 * l2 + x2 * l1 + x3 * l0 + d6};
  uint32_t y[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) s += (uint32_t)a.to_b[4 * i + j] * (p[j] & 0xffff);
    y[i] = s & 0xffff;
  }
  *out = uint2{y[0] | (y[1] << 16), y[2] | (y[3] << 16)};
}

void launch_query_prep(const uint16_t* src, int batch, int K, int party_id, uint16_t* dst,
                       int rows_out, hipStream_t s, int mirror) {
  IRIS_REQUIRE(mirror >= 0 && mirror <= 2, "mirror mode must be 0, 1 or 2");
  IRIS_REQUIRE(K % 800 == 0, "share length must be a multiple of 800");
  IRIS_REQUIRE(batch * kRotations <= rows_out, "batch larger than query_length / 31");
  const QueryPrepConsts c = query_prep_consts(party_id);
  QueryPrepArgs a{};
  a.src = src;
  a.dst = dst;
  a.K = K;
  a.batch = batch;
  a.rows_out = rows_out;
  a.mirror = mirror;
  for (int i = 0; i < 4; i++) a.lag[i] = c.lagrange[i];
  for (int i = 0; i < 16; i++) a.to_b[i] = c.to_b[i];
  const size_t n = (size_t)rows_out * (K / 4);
  query_prep_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
