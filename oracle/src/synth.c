/* Oracle twin of the synthetic workload generator. TEST INFRASTRUCTURE ONLY.
 *
 * Generator definition (shared with the product's GPU generator, which must
 * produce identical bytes):
 *   keystream  = ChaCha20, key = 32 zero bytes, 64-bit nonce, 64-bit counter
 *   plaintext iris  (dataset d, index i): nonce = (d << 8) | 1,
 *       counter i*256 .. : blocks 0..24   -> 200 u64 code words (LE u32 pairs)
 *                          blocks 25..224 -> 6400 u16 u_k; mask pair bits
 *                          (2k, 2k+1) are set iff u_k < 55706 (p = 0.85)
 *   bit flips  (dataset d, index i, n): nonce = (d << 8) | 2, counter i*1024,
 *       partial Fisher-Yates over the 12800 bit positions with u32 draws
 *       (position j swaps with j + (r_j mod (12800 - j))), first n flipped
 *   share randomness (dataset d, index i): nonce = (d << 8) | 3,
 *       counter i*1024: 3200 GR elements (12800 u16) for the code sharing,
 *       then 1600 GR elements (6400 u16) for the trimmed mask sharing.
 * The mask follows the pairing rule of iris_db/iris.rs:244-259 so that the
 * trimmed-mask dot times two equals the full mask dot. */
#include <stdlib.h>
#include <string.h>
#include "../oracle.h"

static const uint32_t ZERO_KEY[8] = {0, 0, 0, 0, 0, 0, 0, 0};

void orc_synth_iris(uint64_t dataset, uint64_t index, uint64_t code[200], uint64_t mask[200]) {
  uint32_t ks[225 * 16];
  orc_chacha_keystream(ZERO_KEY, (dataset << 8) | 1, index * 256, 20, ks, 225 * 16);
  for (int w = 0; w < 200; w++) code[w] = (uint64_t)ks[2 * w] | ((uint64_t)ks[2 * w + 1] << 32);
  memset(mask, 0, 1600);
  const uint16_t *u = (const uint16_t *)(ks + 25 * 16);
  for (int k = 0; k < 6400; k++)
    if (u[k] < 55706) mask[(2 * k) / 64] |= 3ULL << ((2 * k) % 64);
}

void orc_synth_flip(const uint64_t code_in[200], uint64_t dataset, uint64_t index,
                    uint32_t n_flips, uint64_t code_out[200]) {
  if (code_out != code_in) memcpy(code_out, code_in, 1600);
  if (n_flips > 12800) n_flips = 12800;
  uint16_t *perm = malloc(12800 * 2);
  for (int j = 0; j < 12800; j++) perm[j] = (uint16_t)j;
  uint32_t *r = malloc(12800 * 4);
  orc_chacha_keystream(ZERO_KEY, (dataset << 8) | 2, index * 1024, 20, r, 12800);
  for (uint32_t j = 0; j < n_flips; j++) {
    uint32_t k = j + r[j] % (12800 - j);
    uint16_t t = perm[j];
    perm[j] = perm[k];
    perm[k] = t;
    code_out[perm[j] / 64] ^= 1ULL << (perm[j] % 64);
  }
  free(perm);
  free(r);
}

void orc_synth_share_rand(uint64_t dataset, uint64_t index, uint16_t *code_rand,
                          uint16_t *mask_rand) {
  uint32_t *ks = malloc(600 * 16 * 4);
  orc_chacha_keystream(ZERO_KEY, (dataset << 8) | 3, index * 1024, 20, ks, 600 * 16);
  memcpy(code_rand, ks, 12800 * 2);
  memcpy(mask_rand, ks + 6400, 6400 * 2);
  free(ks);
}

void orc_synth_shares(const uint64_t code[200], const uint64_t mask[200], uint64_t dataset,
                      uint64_t index, uint16_t *code_sh, uint16_t *mask_sh) {
  uint16_t *cr = malloc(12800 * 2), *mr = malloc(6400 * 2);
  orc_synth_share_rand(dataset, index, cr, mr);
  orc_encode_iris_code(code, mask, cr, code_sh);
  orc_encode_mask_trimmed(mask, mr, mask_sh);
  free(cr);
  free(mr);
}
