/* Oracle: plaintext iris codes and the share-GEMM arithmetic.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plaintext: iris-mpc-common/src/iris_db/iris.rs:113-135 (from_base64: BE u64
 * then reverse_bits), :143-155 (rotate by 4 bits per column inside 800-bit
 * rows), :534-560 (fractional Hamming distance), :613-633 (all_rotations).
 * Share GEMM: iris-mpc-gpu/src/dot/share_db.rs:45-59 (limbs), :265-296 (sums),
 * :62-112,800-845 (three limb GEMMs), dot/kernel.cu:16-28 (correction + mask).
 * CPU baseline: iris-mpc-cpu/src/protocol/ops.rs:238-347. */
#include <string.h>
#include <stdlib.h>
#include "../oracle.h"

static int b64v(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+') return 62;
  if (c == '/') return 63;
  return -1;
}

int orc_iris_from_base64(const char *s, uint64_t out[200]) {
  uint8_t bytes[1600 + 8];
  size_t nb = 0;
  uint32_t acc = 0;
  int bits = 0;
  for (const char *p = s; *p; p++) {
    if (*p == '=' || *p == '\n' || *p == '\r' || *p == ' ') continue;
    int v = b64v(*p);
    if (v < 0) return -1;
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      if (nb >= sizeof(bytes)) return -2;
      bytes[nb++] = (uint8_t)(acc >> bits);
    }
  }
  if (nb != 1600) return -3;
  for (int w = 0; w < 200; w++) {
    uint64_t be = 0;
    for (int k = 0; k < 8; k++) be = (be << 8) | bytes[8 * w + k];
    uint64_t r = 0;
    for (int k = 0; k < 64; k++) r |= ((be >> k) & 1ULL) << (63 - k);
    out[w] = r;
  }
  return 0;
}

static inline int gb(const uint64_t *a, size_t i) { return (int)((a[i / 64] >> (i % 64)) & 1); }

void orc_iris_rotate_left(uint64_t code[200], int by) {
  uint64_t out[200];
  memset(out, 0, sizeof(out));
  long sh = 4L * by;
  for (size_t base = 0; base < 12800; base += 800)
    for (long p = 0; p < 800; p++) {
      long src = (p + sh) % 800;
      if (src < 0) src += 800;
      if (gb(code, base + (size_t)src)) out[(base + p) / 64] |= 1ULL << ((base + p) % 64);
    }
  memcpy(code, out, sizeof(out));
}

void orc_iris_all_rotations(const uint64_t code[200], uint64_t *out) {
  for (int t = 0; t < ORC_ROTATIONS; t++) {
    memcpy(out + (size_t)t * 200, code, 1600);
    orc_iris_rotate_left(out + (size_t)t * 200, 15 - t);
  }
}

void orc_iris_distance(const uint64_t *c1, const uint64_t *m1, const uint64_t *c2,
                       const uint64_t *m2, uint32_t *hd, uint32_t *md) {
  uint32_t h = 0, m = 0;
  for (int i = 0; i < 200; i++) {
    uint64_t cm = m1[i] & m2[i];
    m += (uint32_t)__builtin_popcountll(cm);
    h += (uint32_t)__builtin_popcountll((c1[i] ^ c2[i]) & cm);
  }
  *hd = h;
  *md = m;
}

int orc_mpc_predicate(uint16_t code_dot, uint16_t mask_dot, uint32_t threshold_a) {
  uint32_t r = (uint32_t)mask_dot * threshold_a - ((uint32_t)code_dot << 16) - 1u;
  return (int)(r >> 31);
}

void orc_mpc_predicate_many(const uint16_t *code_dot, const uint16_t *mask_dot, size_t n,
                            uint32_t threshold_a, uint8_t *out) {
  for (size_t i = 0; i < n; i++) out[i] = (uint8_t)orc_mpc_predicate(code_dot[i], mask_dot[i], threshold_a);
}

/* Plaintext brute force of one query against n DB irises at the MPC
 * comparison's integer predicate: out[d] = OR over the 31 rotations t of
 * orc_mpc_predicate(M - 2H, M, A) with the query's code and mask rotated by
 * rotation t (orc_iris_all_rotations), M = |qm_t & dm_d| and
 * H = |(qc_t ^ dc_d) & qm_t & dm_d| (the masked Hamming distance of
 * iris.rs:534-560; the opened code / mask dots are M - 2H and M). */
void orc_plain_match(const uint64_t qc[200], const uint64_t qm[200], const uint64_t *dc,
                     const uint64_t *dm, size_t n, uint32_t threshold_a, uint8_t *out, int threads) {
  uint64_t *rc = malloc(sizeof(uint64_t) * ORC_ROTATIONS * 200);
  uint64_t *rm = malloc(sizeof(uint64_t) * ORC_ROTATIONS * 200);
  orc_iris_all_rotations(qc, rc);
  orc_iris_all_rotations(qm, rm);
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
  for (long d = 0; d < (long)n; d++) {
    uint8_t hit = 0;
    for (int t = 0; t < ORC_ROTATIONS && !hit; t++) {
      uint32_t h, m;
      orc_iris_distance(rc + (size_t)t * 200, rm + (size_t)t * 200, dc + (size_t)d * 200,
                        dm + (size_t)d * 200, &h, &m);
      hit = (uint8_t)orc_mpc_predicate((uint16_t)(m - 2 * h), (uint16_t)m, threshold_a);
    }
    out[d] = hit;
  }
  free(rc);
  free(rm);
}

void orc_limbs(const uint16_t *v, size_t n, int8_t *l0, int8_t *l1) {
  for (size_t i = 0; i < n; i++) {
    l0[i] = (int8_t)((int)(v[i] & 0xff) - 128);
    l1[i] = (int8_t)((int)(v[i] >> 8) - 128);
  }
}

uint32_t orc_limb_sum(const int8_t *l, size_t n) {
  uint32_t s = 0;
  for (size_t i = 0; i < n; i++) s += (uint32_t)(int32_t)l[i];
  return s;
}

void orc_u16_dot(const uint16_t *q, const uint16_t *db, size_t Qr, size_t N, size_t K,
                 uint16_t *out, int threads) {
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : 1)
  for (long d = 0; d < (long)N; d++) {
    const uint16_t *row = db + (size_t)d * K;
    for (size_t i = 0; i < Qr; i++) {
      const uint16_t *qr = q + i * K;
      uint16_t s = 0;
      for (size_t k = 0; k < K; k++) s = (uint16_t)(s + (uint32_t)qr[k] * row[k]);
      out[i * N + (size_t)d] = s;
    }
  }
}

void orc_limb_dot_corrected(const uint16_t *q, const uint16_t *db, size_t Qr, size_t N,
                            size_t K, uint16_t multiplier, uint16_t *out) {
  int8_t *q0 = malloc(Qr * K), *q1 = malloc(Qr * K), *d0 = malloc(N * K), *d1 = malloc(N * K);
  orc_limbs(q, Qr * K, q0, q1);
  orc_limbs(db, N * K, d0, d1);
  for (size_t i = 0; i < Qr; i++) {
    uint32_t qs0 = orc_limb_sum(q0 + i * K, K), qs1 = orc_limb_sum(q1 + i * K, K);
    for (size_t d = 0; d < N; d++) {
      /* gemm(alpha=1) + gemm(alpha=256, beta=1) x2, wrapping i32 */
      int32_t c00 = 0, c01 = 0, c10 = 0;
      for (size_t k = 0; k < K; k++) {
        c00 += (int32_t)d0[d * K + k] * q0[i * K + k];
        c01 += (int32_t)d0[d * K + k] * q1[i * K + k];
        c10 += (int32_t)d1[d * K + k] * q0[i * K + k];
      }
      uint32_t c = (uint32_t)c00 + ((uint32_t)c01 << 8) + ((uint32_t)c10 << 8);
      uint32_t s0 = orc_limb_sum(d0 + d * K, K) + qs0;
      uint32_t s1 = orc_limb_sum(d1 + d * K, K) + qs1;
      uint16_t r = (uint16_t)(c + (s0 << 7) + ((s0 + s1) << 15));
      out[i * N + d] = (uint16_t)(r * multiplier);
    }
  }
  free(q0); free(q1); free(d0); free(d1);
}

void orc_cpu_rotation_aware_dot(const uint16_t *q_code, const uint16_t *q_mask,
                                const uint16_t *db_code, const uint16_t *db_mask, size_t n,
                                uint16_t *out, int threads) {
  enum { ROW = 800, CROWS = 16, MROWS = 8 };
  /* prerotated row-major query: [row][rotation][800] */
  uint16_t *pc = malloc(sizeof(uint16_t) * CROWS * ORC_ROTATIONS * ROW);
  uint16_t *pm = malloc(sizeof(uint16_t) * MROWS * ORC_ROTATIONS * ROW);
  for (int r = 0; r < CROWS; r++)
    for (int t = 0; t < ORC_ROTATIONS; t++)
      for (int p = 0; p < ROW; p++) {
        int src = (p + 4 * (15 - t)) % ROW;
        if (src < 0) src += ROW;
        pc[((size_t)r * ORC_ROTATIONS + t) * ROW + p] = q_code[r * ROW + src];
        if (r < MROWS) pm[((size_t)r * ORC_ROTATIONS + t) * ROW + p] = q_mask[r * ROW + src];
      }
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
  for (long t = 0; t < (long)n; t++) {
    uint16_t acc_c[ORC_ROTATIONS], acc_m[ORC_ROTATIONS];
    memset(acc_c, 0, sizeof(acc_c));
    memset(acc_m, 0, sizeof(acc_m));
    const uint16_t *tc = db_code + (size_t)t * ORC_CODE_LEN;
    const uint16_t *tm = db_mask + (size_t)t * ORC_MASK_LEN;
    for (int r = 0; r < CROWS; r++)
      for (int rot = 0; rot < ORC_ROTATIONS; rot++) {
        const uint16_t *qr = pc + ((size_t)r * ORC_ROTATIONS + rot) * ROW;
        uint16_t s = 0;
        for (int p = 0; p < ROW; p++) s = (uint16_t)(s + (uint32_t)qr[p] * tc[r * ROW + p]);
        acc_c[rot] = (uint16_t)(acc_c[rot] + s);
      }
    for (int r = 0; r < MROWS; r++)
      for (int rot = 0; rot < ORC_ROTATIONS; rot++) {
        const uint16_t *qr = pm + ((size_t)r * ORC_ROTATIONS + rot) * ROW;
        uint16_t s = 0;
        for (int p = 0; p < ROW; p++) s = (uint16_t)(s + (uint32_t)qr[p] * tm[r * ROW + p]);
        acc_m[rot] = (uint16_t)(acc_m[rot] + s);
      }
    for (int rot = 0; rot < ORC_ROTATIONS; rot++) {
      out[((size_t)t * ORC_ROTATIONS + rot) * 2] = acc_c[rot];
      out[((size_t)t * ORC_ROTATIONS + rot) * 2 + 1] = (uint16_t)(2u * acc_m[rot]);
    }
  }
  free(pc);
  free(pm);
}

void orc_mask_dot_share(uint16_t *dot, size_t n_elems, uint16_t multiplier,
                        const uint32_t seed_own[8], const uint32_t seed_prev[8], uint64_t nonce,
                        uint64_t counter) {
  size_t n_words = (n_elems + 1) / 2;
  n_words = (n_words + 15) / 16 * 16;
  uint32_t *k0 = malloc(n_words * 4), *k1 = malloc(n_words * 4);
  orc_chacha_keystream(seed_own, nonce, counter, 12, k0, n_words);
  orc_chacha_keystream(seed_prev, nonce, counter, 12, k1, n_words);
  const uint16_t *r0 = (const uint16_t *)k0, *r1 = (const uint16_t *)k1; /* LE u16 view */
  for (size_t i = 0; i < n_elems; i++)
    dot[i] = (uint16_t)((uint32_t)dot[i] * multiplier + r0[i] - r1[i]);
  free(k0);
  free(k1);
}

void orc_open_bitmap(const uint64_t *r1, const uint64_t *r2, const uint64_t *r3,
                     size_t n_words, size_t chunk_len, size_t query_len, size_t offset,
                     size_t real_chunk_len, size_t total_db_len, uint64_t *bitmap) {
  for (size_t w = 0; w < n_words; w++) {
    uint64_t v = r1[w] ^ r2[w] ^ r3[w];
    for (int i = 0; i < 64; i++) {
      if (!((v >> i) & 1)) continue;
      size_t e = w * 64 + (size_t)i;
      size_t q = e / chunk_len, d = e % chunk_len;
      if (q >= query_len || d >= real_chunk_len) continue;
      size_t bit = total_db_len * (q / ORC_ROTATIONS) + d + offset;
      bitmap[bit / 64] |= 1ULL << (bit % 64);
    }
  }
}

/* IrisCode::mirrored (iris_db/iris.rs:587-601): bit i moves to
 * old(mirror(new(i))); code bits of the imaginary part (i odd) are inverted,
 * mask bits move unchanged. */
void orc_iris_mirrored(const uint64_t code[200], const uint64_t mask[200], uint64_t out_code[200],
                       uint64_t out_mask[200]) {
  memset(out_code, 0, 1600);
  memset(out_mask, 0, 1600);
  for (size_t i = 0; i < 12800; i++) {
    size_t m = orc_remap_new_to_old(orc_remap_new_to_mirrored(orc_remap_old_to_new(i)));
    uint64_t cb = (code[i / 64] >> (i % 64)) & 1, mb = (mask[i / 64] >> (i % 64)) & 1;
    if (i % 2) cb ^= 1;
    out_code[m / 64] |= cb << (m % 64);
    out_mask[m / 64] |= mb << (m % 64);
  }
}
