// Synthetic iris workloads on the device.
//
// Stream definitions (ChaCha20, all-zero key, 64-bit counter in state words
// 12-13, 64-bit nonce in words 14-15):
//   plaintext iris (dataset d, index i): nonce (d << 8) | 1, counters
//     256 i + 0 .. 24   -> the 200 code words (u64 = little-endian u32 pair),
//     256 i + 25 .. 224 -> 6400 u16 draws u_k; mask bits 2k and 2k + 1 are set
//     iff u_k < 55706 (p = 0.85), the paired layout of iris_db/iris.rs:244-259
//     that makes the trimmed-mask dot times two equal the full mask dot;
//   bit flips (d, i, n): nonce (d << 8) | 2, counters 1024 i ..: 12800 u32
//     draws r_j; a partial Fisher-Yates pass over the 12800 bit positions
//     (position j swaps with j + r_j mod (12800 - j)) flips the first n;
//   sharing randomness (d, i): nonce (d << 8) | 3, counters 1024 i ..: 12800
//     u16 for the code sharing, then 6400 u16 for the trimmed-mask sharing.
// Encoding (GaloisRingIrisCodeShare::encode_iris_code / encode_mask_code,
// iris-mpc-common/src/galois_engine.rs:319-397): coefficient i of the share
// vector ("new" order) takes plaintext bit remap_new_to_old(i)
// (galois_engine.rs:256-264); code value mask - 2 (code & mask), mask value the
// mask bit; every 4 coefficients form a GR(2^16,4) element in basis A, which
// goes to the monomial basis and is shared with degree 1 at the base points
// 1, x, 1 + x: share_p = s + r * p.
#include "synth.hpp"

#include <cstring>
#include <vector>

#include "chacha.hpp"

namespace irismpc {
namespace {

constexpr uint16_t kSynthS[4][4] = {  // basis A -> monomial (gr4_encoding_example.sage:4-11)
    {1, 0, 0, 0}, {58082, 1, 0, 0}, {60579, 25194, 1, 0}, {17325, 51956, 57011, 1}};
constexpr uint32_t kMaskThreshold = 55706;

__host__ __device__ inline ChaChaKey zero_key() {
  ChaChaKey k{};
  return k;
}

__device__ __forceinline__ uint32_t remap_new_to_old(uint32_t i) {
  const uint32_t b = i / 6400, w = i % 6400 / 3200, r1 = i % 3200 / 800, c = i % 800 / 4, r2 = i % 4;
  return 800 * (r2 * 4 + r1) + c * 4 + w * 2 + b;
}

__global__ __launch_bounds__(256) void synth_plain_kernel(uint64_t dataset, uint64_t index0, size_t n,
                                                          uint64_t* codes, uint64_t* masks) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 225) return;
  const size_t iris = t / 225;
  const int blk = (int)(t % 225);
  uint32_t ks[16];
  chacha_block<20>(zero_key(), (dataset << 8) | 1, (index0 + iris) * 256 + blk, ks);
  if (blk < 25) {
    uint64_t* c = codes + iris * kIrisWords + blk * 8;
#pragma unroll
    for (int w = 0; w < 8; w++) c[w] = (uint64_t)ks[2 * w] | ((uint64_t)ks[2 * w + 1] << 32);
  } else {
    uint64_t word = 0;
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const uint32_t u = (ks[k >> 1] >> (16 * (k & 1))) & 0xffffu;
      if (u < kMaskThreshold) word |= 3ull << (2 * k);
    }
    masks[iris * kIrisWords + (blk - 25)] = word;
  }
}

// Thread per ChaCha20 block of sharing randomness: blocks 0..399 of an iris
// cover code elements 8 b .. 8 b + 7, blocks 400..599 trimmed-mask elements.
__global__ __launch_bounds__(256) void synth_shares_kernel(const uint64_t* codes, const uint64_t* masks,
                                                           size_t n, uint64_t rand_dataset,
                                                           uint64_t rand_index0, uint16_t* c0,
                                                           uint16_t* c1, uint16_t* c2, uint16_t* m0,
                                                           uint16_t* m1, uint16_t* m2) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 600) return;
  const size_t iris = t / 600;
  const int blk = (int)(t % 600);
  uint32_t ks[16];
  chacha_block<20>(zero_key(), (rand_dataset << 8) | 3, (rand_index0 + iris) * 1024 + blk, ks);
  const uint64_t* code = codes + iris * kIrisWords;
  const uint64_t* mask = masks + iris * kIrisWords;
  const bool is_code = blk < 400;
  const uint32_t e0 = is_code ? 8 * blk : 8 * (blk - 400);
  const size_t K = is_code ? (size_t)kCodeLen : (size_t)kMaskLen;
  uint16_t* out[3] = {is_code ? c0 : m0, is_code ? c1 : m1, is_code ? c2 : m2};
#pragma unroll
  for (int el = 0; el < 8; el++) {
    const uint32_t e = e0 + el;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t o = remap_new_to_old(4 * e + k);
      const uint32_t mb = (uint32_t)(mask[o >> 6] >> (o & 63)) & 1u;
      const uint32_t cb = (uint32_t)(code[o >> 6] >> (o & 63)) & 1u;
      v[k] = is_code ? mb - 2u * (cb & mb) : mb;
    }
    uint32_t s[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      uint32_t acc = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) acc += (uint32_t)kSynthS[i][j] * v[j];
      s[i] = acc;
    }
    const uint32_t w0 = ks[el * 2], w1 = ks[el * 2 + 1];
    const uint32_t r0 = w0 & 0xffffu, r1 = w0 >> 16, r2 = w1 & 0xffffu, r3 = w1 >> 16;
    // r * x = r3 + (r0 + r3) x + r1 x^2 + r2 x^3  (x^4 = x + 1)
    const uint32_t rx[4] = {r3, r0 + r3, r1, r2};
    const uint32_t sh0[4] = {s[0] + r0, s[1] + r1, s[2] + r2, s[3] + r3};
    const uint32_t sh1[4] = {s[0] + rx[0], s[1] + rx[1], s[2] + rx[2], s[3] + rx[3]};
    const uint32_t sh2[4] = {sh0[0] + rx[0], sh0[1] + rx[1], sh0[2] + rx[2], sh0[3] + rx[3]};
    const uint32_t* sh[3] = {sh0, sh1, sh2};
#pragma unroll
    for (int p = 0; p < 3; p++) {
      uint2* dst = reinterpret_cast<uint2*>(out[p] + iris * K + 4 * (size_t)e);
      *dst = uint2{(sh[p][0] & 0xffffu) | (sh[p][1] << 16), (sh[p][2] & 0xffffu) | (sh[p][3] << 16)};
    }
  }
}

}  // namespace

void synth_iris_host(uint64_t dataset, uint64_t index, uint64_t code[kIrisWords],
                     uint64_t mask[kIrisWords]) {
  uint32_t ks[16];
  for (int blk = 0; blk < 225; blk++) {
    chacha_block<20>(zero_key(), (dataset << 8) | 1, index * 256 + blk, ks);
    if (blk < 25) {
      for (int w = 0; w < 8; w++) code[blk * 8 + w] = (uint64_t)ks[2 * w] | ((uint64_t)ks[2 * w + 1] << 32);
    } else {
      uint64_t word = 0;
      for (int k = 0; k < 32; k++)
        if (((ks[k >> 1] >> (16 * (k & 1))) & 0xffffu) < kMaskThreshold) word |= 3ull << (2 * k);
      mask[blk - 25] = word;
    }
  }
}

void synth_flip_host(uint64_t code[kIrisWords], uint64_t dataset, uint64_t index, uint32_t n_flips) {
  constexpr uint32_t kBits = 12800;
  if (n_flips > kBits) n_flips = kBits;
  std::vector<uint16_t> perm(kBits);
  for (uint32_t j = 0; j < kBits; j++) perm[j] = (uint16_t)j;
  uint32_t ks[16];
  for (uint32_t j = 0; j < n_flips; j++) {
    if (j % 16 == 0) chacha_block<20>(zero_key(), (dataset << 8) | 2, index * 1024 + j / 16, ks);
    const uint32_t k = j + ks[j % 16] % (kBits - j);
    std::swap(perm[j], perm[k]);
    code[perm[j] / 64] ^= 1ull << (perm[j] % 64);
  }
}

void launch_synth_plain(uint64_t dataset, uint64_t index0, size_t n, uint64_t* codes,
                        uint64_t* masks, hipStream_t s) {
  if (n == 0) return;
  synth_plain_kernel<<<(unsigned)ceil_div(n * 225, 256), 256, 0, s>>>(dataset, index0, n, codes, masks);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_synth_shares(const uint64_t* codes, const uint64_t* masks, size_t n,
                         uint64_t rand_dataset, uint64_t rand_index0, uint16_t* const code_out[3],
                         uint16_t* const mask_out[3], hipStream_t s) {
  if (n == 0) return;
  synth_shares_kernel<<<(unsigned)ceil_div(n * 600, 256), 256, 0, s>>>(
      codes, masks, n, rand_dataset, rand_index0, code_out[0], code_out[1], code_out[2], mask_out[0],
      mask_out[1], mask_out[2]);
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
