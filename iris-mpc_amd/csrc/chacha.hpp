// In-register ChaCha block function for gfx950.
//
// Same keystream as iris-mpc-gpu/src/rng/chacha.cu:74-138 (state layout of
// rng/chacha.rs:205-250: constants, key[8], 64-bit block counter in words
// 12-13, 64-bit nonce in words 14-15; 12 rounds for the protocol RNG). The
// reference materialises keystream buffers in HBM with separate fill launches;
// here every consumer computes the block it needs from its counter, so
// randomness costs VALU cycles and no memory traffic.
#pragma once
#include "common.hpp"

namespace irismpc {

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, int n) {
  return __builtin_rotateleft32(x, n);
}

#define IRIS_QR(a, b, c, d)   \
  a += b; d = rotl32(d ^ a, 16); \
  c += d; b = rotl32(b ^ c, 12); \
  a += b; d = rotl32(d ^ a, 8);  \
  c += d; b = rotl32(b ^ c, 7);

// Writes the 16-word block for (key, nonce, counter) into out[].
template <int ROUNDS>
__host__ __device__ __forceinline__ void chacha_block(const ChaChaKey& key, uint64_t nonce,
                                             uint64_t counter, uint32_t out[16]) {
  const uint32_t i0 = 0x61707865u, i1 = 0x3320646eu, i2 = 0x79622d32u, i3 = 0x6b206574u;
  const uint32_t i12 = (uint32_t)counter, i13 = (uint32_t)(counter >> 32);
  const uint32_t i14 = (uint32_t)nonce, i15 = (uint32_t)(nonce >> 32);
  uint32_t x0 = i0, x1 = i1, x2 = i2, x3 = i3;
  uint32_t x4 = key.k[0], x5 = key.k[1], x6 = key.k[2], x7 = key.k[3];
  uint32_t x8 = key.k[4], x9 = key.k[5], x10 = key.k[6], x11 = key.k[7];
  uint32_t x12 = i12, x13 = i13, x14 = i14, x15 = i15;
#pragma unroll
  for (int r = 0; r < ROUNDS; r += 2) {
    IRIS_QR(x0, x4, x8, x12);
    IRIS_QR(x1, x5, x9, x13);
    IRIS_QR(x2, x6, x10, x14);
    IRIS_QR(x3, x7, x11, x15);
    IRIS_QR(x0, x5, x10, x15);
    IRIS_QR(x1, x6, x11, x12);
    IRIS_QR(x2, x7, x8, x13);
    IRIS_QR(x3, x4, x9, x14);
  }
  out[0] = x0 + i0; out[1] = x1 + i1; out[2] = x2 + i2; out[3] = x3 + i3;
  out[4] = x4 + key.k[0]; out[5] = x5 + key.k[1]; out[6] = x6 + key.k[2]; out[7] = x7 + key.k[3];
  out[8] = x8 + key.k[4]; out[9] = x9 + key.k[5]; out[10] = x10 + key.k[6]; out[11] = x11 + key.k[7];
  out[12] = x12 + i12; out[13] = x13 + i13; out[14] = x14 + i14; out[15] = x15 + i15;
}

// The same block in pieces, for consumers that spread one block's rounds over
// other latency-bound work: chacha_start, ROUNDS / 2 double rounds (each a
// column and a diagonal round), chacha_finish (which adds the input words
// back in place).
__device__ __forceinline__ void chacha_start(const ChaChaKey& key, uint64_t nonce, uint64_t counter,
                                             uint32_t x[16]) {
  x[0] = 0x61707865u; x[1] = 0x3320646eu; x[2] = 0x79622d32u; x[3] = 0x6b206574u;
#pragma unroll
  for (int i = 0; i < 8; i++) x[4 + i] = key.k[i];
  x[12] = (uint32_t)counter; x[13] = (uint32_t)(counter >> 32);
  x[14] = (uint32_t)nonce; x[15] = (uint32_t)(nonce >> 32);
}

// A double round is a column round, then a diagonal round.
__device__ __forceinline__ void chacha_column_round(uint32_t x[16]) {
  IRIS_QR(x[0], x[4], x[8], x[12]);
  IRIS_QR(x[1], x[5], x[9], x[13]);
  IRIS_QR(x[2], x[6], x[10], x[14]);
  IRIS_QR(x[3], x[7], x[11], x[15]);
}

__device__ __forceinline__ void chacha_diagonal_round(uint32_t x[16]) {
  IRIS_QR(x[0], x[5], x[10], x[15]);
  IRIS_QR(x[1], x[6], x[11], x[12]);
  IRIS_QR(x[2], x[7], x[8], x[13]);
  IRIS_QR(x[3], x[4], x[9], x[14]);
}

__device__ __forceinline__ void chacha_finish(const ChaChaKey& key, uint64_t nonce, uint64_t counter,
                                              uint32_t x[16]) {
  x[0] += 0x61707865u; x[1] += 0x3320646eu; x[2] += 0x79622d32u; x[3] += 0x6b206574u;
#pragma unroll
  for (int i = 0; i < 8; i++) x[4 + i] += key.k[i];
  x[12] += (uint32_t)counter; x[13] += (uint32_t)(counter >> 32);
  x[14] += (uint32_t)nonce; x[15] += (uint32_t)(nonce >> 32);
}

#undef IRIS_QR

}  // namespace irismpc
