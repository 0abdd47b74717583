// SHA-256 (FIPS 180-4), HMAC (RFC 2104) and HKDF (RFC 5869) for derive_seed.
#include "host_crypto.hpp"

#include <array>
#include <cstring>
#include <vector>

namespace irismpc {
namespace {
constexpr std::array<uint32_t, 64> kK = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void compress(uint32_t h[8], const uint8_t* blk) {
  uint32_t w[64];
  for (int i = 0; i < 16; i++)
    w[i] = (uint32_t)blk[4 * i] << 24 | (uint32_t)blk[4 * i + 1] << 16 |
           (uint32_t)blk[4 * i + 2] << 8 | blk[4 * i + 3];
  for (int i = 16; i < 64; i++)
    w[i] = w[i - 16] + (rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3)) + w[i - 7] +
           (rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10));
  uint32_t v[8];
  std::memcpy(v, h, 32);
  for (int i = 0; i < 64; i++) {
    const uint32_t t1 = v[7] + (rotr(v[4], 6) ^ rotr(v[4], 11) ^ rotr(v[4], 25)) +
                        ((v[4] & v[5]) ^ (~v[4] & v[6])) + kK[i] + w[i];
    const uint32_t t2 = (rotr(v[0], 2) ^ rotr(v[0], 13) ^ rotr(v[0], 22)) +
                        ((v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]));
    for (int j = 7; j > 0; j--) v[j] = v[j - 1];
    v[4] += t1;
    v[0] = t1 + t2;
  }
  for (int i = 0; i < 8; i++) h[i] += v[i];
}
}  // namespace

void sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
  uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                   0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  std::vector<uint8_t> m(msg, msg + len);
  m.push_back(0x80);
  while (m.size() % 64 != 56) m.push_back(0);
  const uint64_t bits = (uint64_t)len * 8;
  for (int i = 7; i >= 0; i--) m.push_back((uint8_t)(bits >> (8 * i)));
  for (size_t off = 0; off < m.size(); off += 64) compress(h, m.data() + off);
  for (int i = 0; i < 8; i++)
    for (int b = 0; b < 4; b++) out[4 * i + b] = (uint8_t)(h[i] >> (24 - 8 * b));
}

void hmac_sha256(const uint8_t* key, size_t klen, const uint8_t* msg, size_t mlen, uint8_t out[32]) {
  uint8_t k[64] = {0};
  if (klen > 64)
    sha256(key, klen, k);
  else
    std::memcpy(k, key, klen);
  std::vector<uint8_t> inner(64 + mlen), outer(64 + 32);
  for (int i = 0; i < 64; i++) {
    inner[i] = k[i] ^ 0x36;
    outer[i] = k[i] ^ 0x5c;
  }
  if (mlen) std::memcpy(inner.data() + 64, msg, mlen);
  sha256(inner.data(), inner.size(), outer.data() + 64);
  sha256(outer.data(), outer.size(), out);
}

void hkdf_sha256_32(const uint8_t* salt, size_t slen, const uint8_t* ikm, size_t ilen,
                    const uint8_t* info, size_t infolen, uint8_t out[32]) {
  uint8_t prk[32];
  hmac_sha256(salt, slen, ikm, ilen, prk);
  std::vector<uint8_t> t(info, info + infolen);
  t.push_back(1);
  hmac_sha256(prk, 32, t.data(), t.size(), out);
}

void derive_seed(const uint32_t seed[8], uint64_t nonce, uint32_t out[8]) {
  static const uint8_t kSalt[32] = {0x11, 0x1a, 0x1a, 0x93, 0x51, 0x8f, 0x67, 0x0e, 0x9b, 0xb0, 0xc2,
                                    0xc6, 0x88, 0x88, 0xe2, 0xbe, 0xb9, 0x40, 0x6d, 0x4c, 0x4e, 0xd5,
                                    0x71, 0xdc, 0x77, 0xb8, 0x01, 0xe6, 0x76, 0xae, 0x30, 0x91};
  uint8_t ikm[32], info[8], okm[32];
  for (int i = 0; i < 8; i++)
    for (int b = 0; b < 4; b++) ikm[4 * i + b] = (uint8_t)(seed[i] >> (8 * b));
  for (int i = 0; i < 8; i++) info[i] = (uint8_t)(nonce >> (56 - 8 * i));
  hkdf_sha256_32(kSalt, 32, ikm, 32, info, 8, okm);
  for (int i = 0; i < 8; i++)
    out[i] = (uint32_t)okm[4 * i] | (uint32_t)okm[4 * i + 1] << 8 | (uint32_t)okm[4 * i + 2] << 16 |
             (uint32_t)okm[4 * i + 3] << 24;
}

}  // namespace irismpc
