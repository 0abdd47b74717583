/* Oracle: ChaCha block function and SHA-256 / HMAC / HKDF.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * ChaCha follows the DJB construction (RFC 8439 s2.1-2.3) with the state layout
 * used by iris-mpc-gpu/src/rng/chacha.rs:205-250 (64-bit counter in words
 * 12-13, 64-bit nonce in words 14-15) and the double-round schedule of
 * iris-mpc-gpu/src/rng/chacha.cu:108-118 (rounds = 12 there).
 * SHA-256 follows FIPS 180-4; HMAC RFC 2104; HKDF RFC 5869, as used by
 * ring 0.17 in iris-mpc-gpu/src/server/actor.rs:2914-2925. */
#include <string.h>
#include "../oracle.h"

#define ROTL(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define QR(a, b, c, d)                                                          \
  a += b; d ^= a; d = ROTL(d, 16);                                              \
  c += d; b ^= c; b = ROTL(b, 12);                                              \
  a += b; d ^= a; d = ROTL(d, 8);                                               \
  c += d; b ^= c; b = ROTL(b, 7);

void orc_chacha_block(const uint32_t in[16], uint32_t out[16], int rounds) {
  uint32_t x[16];
  memcpy(x, in, sizeof(x));
  for (int i = 0; i < rounds; i += 2) {
    QR(x[0], x[4], x[8], x[12]);
    QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]);
    QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]);
    QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]);
    QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

/* Threads for the keystream and the protocol's element loops (1 = serial).
 * Only the order of the work changes with it, never a value. */
int orc_threads = 1;

void orc_set_threads(int n) { orc_threads = n > 0 ? n : 1; }

void orc_chacha_keystream(const uint32_t key[8], uint64_t nonce, uint64_t counter, int rounds,
                          uint32_t *out, size_t n_words) {
  const size_t n_blocks = (n_words + 15) / 16;
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1 && n_blocks >= 4096)
  for (size_t b = 0; b < n_blocks; b++) {
    uint32_t st[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
    memcpy(st + 4, key, 32);
    st[12] = (uint32_t)(counter + b);
    st[13] = (uint32_t)((counter + b) >> 32);
    st[14] = (uint32_t)nonce;
    st[15] = (uint32_t)(nonce >> 32);
    uint32_t blk[16];
    orc_chacha_block(st, blk, rounds);
    const size_t done = b * 16;
    const size_t take = n_words - done < 16 ? n_words - done : 16;
    memcpy(out + done, blk, take * 4);
  }
}

/* ---- SHA-256 ---- */
static const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4,
    0xab1c5ed5, 0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe,
    0x9bdc06a7, 0xc19bf174, 0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f,
    0x4a7484aa, 0x5cb0a9dc, 0x76f988da, 0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7,
    0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967, 0x27b70a85, 0x2e1b2138, 0x4d2c6dfc,
    0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85, 0xa2bfe8a1, 0xa81a664b,
    0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070, 0x19a4c116,
    0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7,
    0xc67178f2};
#define RR(x, n) (((x) >> (n)) | ((x) << (32 - (n))))

typedef struct {
  uint32_t h[8];
  uint8_t buf[64];
  size_t buflen;
  uint64_t total;
} sha_ctx;

static void sha_compress(uint32_t h[8], const uint8_t p[64]) {
  uint32_t w[64];
  for (int i = 0; i < 16; i++)
    w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) |
           ((uint32_t)p[4 * i + 2] << 8) | p[4 * i + 3];
  for (int i = 16; i < 64; i++) {
    uint32_t s0 = RR(w[i - 15], 7) ^ RR(w[i - 15], 18) ^ (w[i - 15] >> 3);
    uint32_t s1 = RR(w[i - 2], 17) ^ RR(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
  for (int i = 0; i < 64; i++) {
    uint32_t S1 = RR(e, 6) ^ RR(e, 11) ^ RR(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = hh + S1 + ch + K256[i] + w[i];
    uint32_t S0 = RR(a, 2) ^ RR(a, 13) ^ RR(a, 22);
    uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    uint32_t t2 = S0 + mj;
    hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

static void sha_init(sha_ctx *c) {
  static const uint32_t iv[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  memcpy(c->h, iv, 32);
  c->buflen = 0;
  c->total = 0;
}

static void sha_update(sha_ctx *c, const uint8_t *m, size_t n) {
  c->total += n;
  while (n > 0) {
    size_t take = 64 - c->buflen < n ? 64 - c->buflen : n;
    memcpy(c->buf + c->buflen, m, take);
    c->buflen += take;
    m += take;
    n -= take;
    if (c->buflen == 64) {
      sha_compress(c->h, c->buf);
      c->buflen = 0;
    }
  }
}

static void sha_final(sha_ctx *c, uint8_t out[32]) {
  uint64_t bits = c->total * 8;
  uint8_t pad = 0x80;
  sha_update(c, &pad, 1);
  uint8_t z = 0;
  while (c->buflen != 56) sha_update(c, &z, 1);
  uint8_t len[8];
  for (int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (56 - 8 * i));
  sha_update(c, len, 8);
  for (int i = 0; i < 8; i++) {
    out[4 * i] = (uint8_t)(c->h[i] >> 24);
    out[4 * i + 1] = (uint8_t)(c->h[i] >> 16);
    out[4 * i + 2] = (uint8_t)(c->h[i] >> 8);
    out[4 * i + 3] = (uint8_t)c->h[i];
  }
}

void orc_sha256(const uint8_t *msg, size_t len, uint8_t out[32]) {
  sha_ctx c;
  sha_init(&c);
  sha_update(&c, msg, len);
  sha_final(&c, out);
}

void orc_hmac_sha256(const uint8_t *key, size_t klen, const uint8_t *msg, size_t mlen,
                     uint8_t out[32]) {
  uint8_t k0[64] = {0};
  if (klen > 64)
    orc_sha256(key, klen, k0);
  else
    memcpy(k0, key, klen);
  uint8_t ipad[64], opad[64];
  for (int i = 0; i < 64; i++) {
    ipad[i] = k0[i] ^ 0x36;
    opad[i] = k0[i] ^ 0x5c;
  }
  uint8_t inner[32];
  sha_ctx c;
  sha_init(&c);
  sha_update(&c, ipad, 64);
  sha_update(&c, msg, mlen);
  sha_final(&c, inner);
  sha_init(&c);
  sha_update(&c, opad, 64);
  sha_update(&c, inner, 32);
  sha_final(&c, out);
}

void orc_hkdf_sha256(const uint8_t *salt, size_t slen, const uint8_t *ikm, size_t ilen,
                     const uint8_t *info, size_t infolen, uint8_t *okm, size_t okmlen) {
  uint8_t prk[32];
  orc_hmac_sha256(salt, slen, ikm, ilen, prk);
  uint8_t t[32];
  size_t tlen = 0, done = 0;
  uint8_t ctr = 1;
  uint8_t msg[32 + 256 + 1];
  while (done < okmlen) {
    size_t m = 0;
    memcpy(msg, t, tlen);
    m += tlen;
    memcpy(msg + m, info, infolen);
    m += infolen;
    msg[m++] = ctr++;
    orc_hmac_sha256(prk, 32, msg, m, t);
    tlen = 32;
    size_t take = okmlen - done < 32 ? okmlen - done : 32;
    memcpy(okm + done, t, take);
    done += take;
  }
}

/* actor.rs:116 KDF_SALT (hex) */
static const char *KDF_SALT_HEX = "111a1a93518f670e9bb0c2c68888e2beb9406d4c4ed571dc77b801e676ae3091";

static int hexval(char c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  return c - 'A' + 10;
}

void orc_derive_seed(const uint32_t seed[8], uint64_t nonce, uint32_t out[8]) {
  uint8_t salt[32];
  for (int i = 0; i < 32; i++)
    salt[i] = (uint8_t)(hexval(KDF_SALT_HEX[2 * i]) * 16 + hexval(KDF_SALT_HEX[2 * i + 1]));
  uint8_t ikm[32];
  for (int i = 0; i < 8; i++)
    for (int b = 0; b < 4; b++) ikm[4 * i + b] = (uint8_t)(seed[i] >> (8 * b)); /* LE cast */
  uint8_t info[8];
  for (int i = 0; i < 8; i++) info[i] = (uint8_t)(nonce >> (56 - 8 * i)); /* usize BE */
  uint8_t okm[32];
  orc_hkdf_sha256(salt, 32, ikm, 32, info, 8, okm, 32);
  for (int i = 0; i < 8; i++)
    out[i] = (uint32_t)okm[4 * i] | ((uint32_t)okm[4 * i + 1] << 8) |
             ((uint32_t)okm[4 * i + 2] << 16) | ((uint32_t)okm[4 * i + 3] << 24);
}
