/* Oracle: the 3-party replicated-share threshold protocol, all parties
 * simulated in one process. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates iris-mpc-gpu/src/threshold_ring/protocol.rs:
 *   compare_threshold_masked_many :2519-2546
 *   lift_mpc                      :2041-2074 (transpose :1673-1709,
 *                                  lift_split :1787-1829)
 *   binary_add_3_get_two_carries  :2137-2241
 *   bit_inject_party_{0,1,2}      :1302-1629
 *   lift_mul_sub                  :1831-1872
 *   extract_msb                   :2243-2258, binary_add_3_get_msb :2262-2359
 *   packed_and_many_pre/and_many_pre :610-710, send_receive_* :894-987,
 *   chacha1/2 OTP                 :784-892
 * and the kernels of threshold_ring/cuda/kernel.cu (and_pre_inner :39-42,
 * transpose16x64/32x64 :63-107, split_inner :205-235, lift_mul_sub :181-195,
 * packed_bit_inject_party_* :526-629).
 *
 * Every exchanged message is OTP-encrypted with the sender's chacha1 stream and
 * decrypted with the receiver's chacha2 stream exactly as the reference does,
 * so a counter-schedule mistake corrupts the opened result. Party p holds
 * replicated shares (a = x_p, b = x_{p-1}); it sends to next = p+1 and
 * receives from prev = p-1 (protocol.rs:498-499). */
#include <stdlib.h>
#include <string.h>
#include "../oracle.h"

extern int orc_threads;
/* element loops split over orc_threads when long (order of work only) */
#define PAR_FOR(n) _Pragma("omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1)")

typedef struct {
  uint32_t s1[8], s2[8]; /* chacha1 = own seed, chacha2 = prev party's seed */
  uint64_t nonce, ctr1, ctr2;
  int id;
} party_t;

#define NEXT(p) (((p) + 1) % 3)
#define PREV(p) (((p) + 2) % 3)
#define CEIL_DIV(a, b) (((a) + (b) - 1) / (b))

static uint32_t *ks_alloc(const uint32_t seed[8], uint64_t nonce, uint64_t *ctr, size_t n_u32) {
  uint32_t *k = malloc(n_u32 * 4 + 4);
  orc_chacha_keystream(seed, nonce, *ctr, 12, k, n_u32);
  *ctr += n_u32 / 16;
  return k;
}

/* ChaChaCudaCorrRng::fill_rng_into: F(s1) ^ F(s2); both counters advance. */
static uint64_t *corr_rand_u64(party_t *P, size_t n_u64) {
  uint32_t *k1 = ks_alloc(P->s1, P->nonce, &P->ctr1, 2 * n_u64);
  uint32_t *k2 = ks_alloc(P->s2, P->nonce, &P->ctr2, 2 * n_u64);
  uint64_t *r = malloc(n_u64 * 8 + 8);
  PAR_FOR(n_u64)
  for (size_t i = 0; i < n_u64; i++)
    r[i] = (uint64_t)(k1[2 * i] ^ k2[2 * i]) | ((uint64_t)(k1[2 * i + 1] ^ k2[2 * i + 1]) << 32);
  free(k1);
  free(k2);
  return r;
}

/* and_pre_inner with correlated randomness drawn for `alloc_u64` words. */
static void and_pre(party_t *P, const uint64_t *xa, const uint64_t *xb, const uint64_t *ya,
                    const uint64_t *yb, uint64_t *res_a, size_t n, size_t alloc_u64) {
  uint64_t *r = corr_rand_u64(P, alloc_u64);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) res_a[i] = (xa[i] & ya[i]) ^ (xb[i] & ya[i]) ^ (xa[i] & yb[i]) ^ r[i];
  free(r);
}

/* send_receive_view_with_offset: send a[0..n) to next, receive into b[0..n). */
static void send_recv_u64(party_t P[3], uint64_t *const a[3], uint64_t *const b[3], size_t n) {
  size_t ks = CEIL_DIV(n, 8);
  uint64_t *msg[3];
  for (int p = 0; p < 3; p++) {
    uint32_t *k = ks_alloc(P[p].s1, P[p].nonce, &P[p].ctr1, 16 * ks);
    msg[p] = malloc(n * 8 + 8);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++)
      msg[p][i] = a[p][i] ^ ((uint64_t)k[2 * i] | ((uint64_t)k[2 * i + 1] << 32));
    free(k);
  }
  for (int p = 0; p < 3; p++) {
    int q = NEXT(p);
    uint32_t *k = ks_alloc(P[q].s2, P[q].nonce, &P[q].ctr2, 16 * ks);
    PAR_FOR(n)
    for (size_t i = 0; i < n; i++)
      b[q][i] = msg[p][i] ^ ((uint64_t)k[2 * i] | ((uint64_t)k[2 * i + 1] << 32));
    free(k);
    free(msg[p]);
  }
}

static void xor_into(uint64_t *d, const uint64_t *s, size_t n) {
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) d[i] ^= s[i];
}

/* Plane j, word i, bit k = bit j of element 64 i + k (kernel.cu:63-107). */
static void transpose_u16(const uint16_t *in, size_t W, int planes, uint64_t *out) {
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1 && W >= 4096)
  for (size_t i = 0; i < W; i++) {
    for (int j = 0; j < planes; j++) out[(size_t)j * W + i] = 0;
    for (int k = 0; k < 64; k++) {
      uint16_t v = in[i * 64 + (size_t)k];
      for (int j = 0; j < planes; j++) out[(size_t)j * W + i] |= (uint64_t)((v >> j) & 1) << k;
    }
  }
}

static void transpose_u32(const uint32_t *in, size_t W, int planes, uint64_t *out) {
#pragma omp parallel for schedule(static) num_threads(orc_threads) if (orc_threads > 1 && W >= 4096)
  for (size_t i = 0; i < W; i++) {
    for (int j = 0; j < planes; j++) out[(size_t)j * W + i] = 0;
    for (int k = 0; k < 64; k++) {
      uint32_t v = in[i * 64 + (size_t)k];
      for (int j = 0; j < planes; j++) out[(size_t)j * W + i] |= (uint64_t)((v >> j) & 1) << k;
    }
  }
}

/* split_inner (kernel.cu:205-235): x1 := (a, b) -> three replicated sharings
 * of the three additive summands. */
static void split(int id, uint64_t *x1a, uint64_t *x1b, uint64_t *x2a, uint64_t *x2b,
                  uint64_t *x3a, uint64_t *x3b, size_t n) {
  for (size_t i = 0; i < n; i++) {
    uint64_t ta = x1a[i], tb = x1b[i];
    switch (id) {
      case 0: x1a[i] = ta; x1b[i] = 0; x2a[i] = 0; x2b[i] = 0; x3a[i] = 0; x3b[i] = tb; break;
      case 1: x1a[i] = 0; x1b[i] = tb; x2a[i] = ta; x2b[i] = 0; x3a[i] = 0; x3b[i] = 0; break;
      default: x1a[i] = 0; x1b[i] = 0; x2a[i] = 0; x2b[i] = tb; x3a[i] = ta; x3b[i] = 0; break;
    }
  }
}

typedef struct {
  uint64_t *a, *b;
} rs_t;

static rs_t rs_new(size_t n) {
  rs_t r = {calloc(n + 1, 8), calloc(n + 1, 8)};
  return r;
}
static void rs_free(rs_t *r) { free(r->a); free(r->b); }

/* binary_add_3_get_two_carries (protocol.rs:2137-2241), K = 16.
 * x1,x2,x3: 16W words each; out c: 2W words (c1 = bit 16, c2 = bit 17). */
static void add3_two_carries(party_t P[3], rs_t x1[3], rs_t x2[3], rs_t x3[3], rs_t c[3],
                             size_t W) {
  const size_t K = 16;
  rs_t s[3], carry[3];
  uint64_t *sa[3], *sb[3];
  for (int p = 0; p < 3; p++) {
    s[p] = rs_new((K - 1) * W);
    carry[p] = rs_new(K * W);
    xor_into(x2[p].a, x3[p].a, K * W); /* x2x3 */
    xor_into(x2[p].b, x3[p].b, K * W);
    for (size_t i = 0; i < (K - 1) * W; i++) {
      s[p].a[i] = x1[p].a[W + i] ^ x2[p].a[W + i];
      s[p].b[i] = x1[p].b[W + i] ^ x2[p].b[W + i];
    }
    xor_into(x1[p].a, x3[p].a, K * W); /* x1x3 */
    xor_into(x1[p].b, x3[p].b, K * W);
    and_pre(&P[p], x1[p].a, x1[p].b, x2[p].a, x2[p].b, carry[p].a, K * W,
            CEIL_DIV(K * W, 8) * 8);
  }
  {
    uint64_t *A[3] = {carry[0].a, carry[1].a, carry[2].a}, *B[3] = {carry[0].b, carry[1].b, carry[2].b};
    send_recv_u64(P, A, B, K * W);
  }
  for (int p = 0; p < 3; p++) {
    xor_into(carry[p].a, x3[p].a, K * W);
    xor_into(carry[p].b, x3[p].b, K * W);
  }
  /* ripple: a = s (planes <-> bit positions 1..15), b = carry (0..15 -> 1..16) */
  for (int p = 0; p < 3; p++)
    and_pre(&P[p], s[p].a, s[p].b, carry[p].a, carry[p].b, c[p].a, W, CEIL_DIV(W, 8) * 8);
  {
    uint64_t *A[3] = {c[0].a, c[1].a, c[2].a}, *B[3] = {c[0].b, c[1].b, c[2].b};
    send_recv_u64(P, A, B, W);
  }
  for (size_t k = 1; k < K - 1; k++) {
    for (int p = 0; p < 3; p++) {
      uint64_t *ak_a = s[p].a + k * W, *ak_b = s[p].b + k * W;
      uint64_t *bk_a = carry[p].a + k * W, *bk_b = carry[p].b + k * W;
      xor_into(ak_a, c[p].a, W);
      xor_into(ak_b, c[p].b, W);
      xor_into(bk_a, c[p].a, W);
      xor_into(bk_b, c[p].b, W);
      and_pre(&P[p], ak_a, ak_b, bk_a, bk_b, s[p].a /* tmp slot a[0] */, W, CEIL_DIV(W, 8) * 8);
      sa[p] = s[p].a;
      sb[p] = s[p].b;
    }
    send_recv_u64(P, sa, sb, W);
    for (int p = 0; p < 3; p++) {
      xor_into(c[p].a, s[p].a, W);
      xor_into(c[p].b, s[p].b, W);
    }
  }
  for (int p = 0; p < 3; p++) {
    uint64_t *b15a = carry[p].a + (K - 1) * W, *b15b = carry[p].b + (K - 1) * W;
    and_pre(&P[p], b15a, b15b, c[p].a, c[p].b, c[p].a + W, W, CEIL_DIV(W, 8) * 8);
    xor_into(c[p].a, b15a, W);
    xor_into(c[p].b, b15b, W);
  }
  {
    uint64_t *A[3] = {c[0].a + W, c[1].a + W, c[2].a + W}, *B[3] = {c[0].b + W, c[1].b + W, c[2].b + W};
    send_recv_u64(P, A, B, W);
  }
  for (int p = 0; p < 3; p++) {
    rs_free(&s[p]);
    rs_free(&carry[p]);
  }
}

static inline uint16_t bit_of(const uint64_t *w, size_t i) { return (uint16_t)((w[i / 64] >> (i % 64)) & 1); }

/* u16 OTP (chacha1_encrypt_u16 / chacha2_decrypt_u16, protocol.rs:784-810). */
static uint16_t *otp_u16(const uint32_t seed[8], uint64_t nonce, uint64_t *ctr, const uint16_t *in,
                         size_t n) {
  uint32_t *k = ks_alloc(seed, nonce, ctr, n / 2);
  const uint16_t *k16 = (const uint16_t *)k;
  uint16_t *o = malloc(n * 2 + 2);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) o[i] = in[i] ^ k16[i];
  free(k);
  return o;
}

/* bit_inject (protocol.rs:1302-1643): c (2W words) -> u16 shares out (128W). */
static void bit_inject(party_t P[3], rs_t c[3], uint16_t *out_a[3], uint16_t *out_b[3], size_t W,
                       size_t alloc) {
  const size_t n = 128 * W;
  const size_t rn = 64 * alloc; /* u32 words per rand buffer */
  /* party 0 */
  uint32_t *r01_0 = ks_alloc(P[0].s1, P[0].nonce, &P[0].ctr1, rn);
  uint32_t *r02_0 = ks_alloc(P[0].s2, P[0].nonce, &P[0].ctr2, rn);
  const uint16_t *R01_0 = (const uint16_t *)r01_0, *R02_0 = (const uint16_t *)r02_0;
  uint16_t *y = malloc(n * 2);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) y[i] = (uint16_t)(bit_of(c[0].b, i) * R01_0[i] - R02_0[i]);
  uint16_t *y_enc = otp_u16(P[0].s1, P[0].nonce, &P[0].ctr1, y, n);
  /* party 1 */
  uint32_t *r01_1 = ks_alloc(P[1].s2, P[1].nonce, &P[1].ctr2, rn);
  uint32_t *r12_1 = ks_alloc(P[1].s1, P[1].nonce, &P[1].ctr1, rn);
  const uint16_t *R01_1 = (const uint16_t *)r01_1, *R12_1 = (const uint16_t *)r12_1;
  uint16_t *x = malloc(n * 2);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++)
    x[i] = (uint16_t)((bit_of(c[1].a, i) ^ bit_of(c[1].b, i)) - R01_1[i]);
  uint16_t *x_enc = otp_u16(P[1].s1, P[1].nonce, &P[1].ctr1, x, n);
  /* party 2: receives x first */
  uint32_t *r12_2 = ks_alloc(P[2].s2, P[2].nonce, &P[2].ctr2, rn);
  uint32_t *r02_2 = ks_alloc(P[2].s1, P[2].nonce, &P[2].ctr1, rn);
  const uint16_t *R12_2 = (const uint16_t *)r12_2, *R02_2 = (const uint16_t *)r02_2;
  uint16_t *x2 = otp_u16(P[2].s2, P[2].nonce, &P[2].ctr2, x_enc, n);
  uint16_t *z = malloc(n * 2);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) z[i] = (uint16_t)(bit_of(c[2].a, i) * x2[i] - R12_2[i]);
  uint16_t *z_enc = otp_u16(P[2].s1, P[2].nonce, &P[2].ctr1, z, n);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) {
    out_a[2][i] = (uint16_t)(-(((uint32_t)R02_2[i] + z[i]) << 1) + bit_of(c[2].a, i));
    out_b[2][i] = (uint16_t)(x2[i] - ((uint32_t)R12_2[i] << 1));
  }
  /* party 0 receives z */
  uint16_t *z0 = otp_u16(P[0].s2, P[0].nonce, &P[0].ctr2, z_enc, n);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) {
    out_a[0][i] = (uint16_t)(R01_0[i] - ((uint32_t)y[i] << 1));
    out_b[0][i] = (uint16_t)(-(((uint32_t)R02_0[i] + z0[i]) << 1) + bit_of(c[0].b, i));
  }
  /* party 1 receives y */
  uint16_t *y1 = otp_u16(P[1].s2, P[1].nonce, &P[1].ctr2, y_enc, n);
  PAR_FOR(n)
  for (size_t i = 0; i < n; i++) {
    out_a[1][i] = (uint16_t)(x[i] - ((uint32_t)R12_1[i] << 1));
    out_b[1][i] = (uint16_t)(R01_1[i] - ((uint32_t)y1[i] << 1));
  }
  free(r01_0); free(r02_0); free(y); free(y_enc);
  free(r01_1); free(r12_1); free(x); free(x_enc);
  free(r12_2); free(r02_2); free(x2); free(z); free(z_enc);
  free(z0); free(y1);
}

static void init_parties(party_t P[3], const uint32_t *seeds_own, const uint32_t *seeds_prev,
                         uint64_t nonce, uint64_t counter) {
  for (int p = 0; p < 3; p++) {
    memcpy(P[p].s1, seeds_own + 8 * p, 32);
    memcpy(P[p].s2, seeds_prev + 8 * p, 32);
    P[p].nonce = nonce;
    P[p].ctr1 = P[p].ctr2 = counter;
    P[p].id = p;
  }
}

/* lift_mpc: mask u16 shares -> lifted u32 (uncorrected) + u16 corrections. */
static void lift_mpc(party_t P[3], const uint16_t *const ma[3], const uint16_t *const mb[3],
                     size_t L, size_t alloc, uint32_t *lift_a[3], uint32_t *lift_b[3],
                     uint16_t *corr_a[3], uint16_t *corr_b[3]) {
  const size_t W = L / 64, K = 16;
  rs_t x1[3], x2[3], x3[3], c[3];
  for (int p = 0; p < 3; p++) {
    x1[p] = rs_new(K * W); x2[p] = rs_new(K * W); x3[p] = rs_new(K * W); c[p] = rs_new(2 * W);
    transpose_u16(ma[p], W, 16, x1[p].a);
    transpose_u16(mb[p], W, 16, x1[p].b);
    for (size_t i = 0; i < L; i++) {
      lift_a[p][i] = ma[p][i];
      lift_b[p][i] = mb[p][i];
    }
    split(p, x1[p].a, x1[p].b, x2[p].a, x2[p].b, x3[p].a, x3[p].b, K * W);
  }
  add3_two_carries(P, x1, x2, x3, c, W);
  bit_inject(P, c, corr_a, corr_b, W, alloc);
  for (int p = 0; p < 3; p++) {
    rs_free(&x1[p]); rs_free(&x2[p]); rs_free(&x3[p]); rs_free(&c[p]);
  }
}

uint64_t orc_lift(const uint16_t *const mask_a[3], const uint16_t *const mask_b[3], size_t L,
                  size_t alloc_words, const uint32_t *seeds_own, const uint32_t *seeds_prev,
                  uint64_t nonce, uint64_t counter, uint32_t *const out_a[3],
                  uint32_t *const out_b[3]) {
  party_t P[3];
  init_parties(P, seeds_own, seeds_prev, nonce, counter);
  uint16_t *ca[3], *cb[3];
  for (int p = 0; p < 3; p++) {
    ca[p] = malloc(2 * L * 2);
    cb[p] = malloc(2 * L * 2);
  }
  lift_mpc(P, mask_a, mask_b, L, alloc_words, (uint32_t **)out_a, (uint32_t **)out_b, ca, cb);
  for (int p = 0; p < 3; p++) {
    for (size_t i = 0; i < L; i++) {
      out_a[p][i] -= ((uint32_t)ca[p][i] << 16) + ((uint32_t)ca[p][i + L] << 17);
      out_b[p][i] -= ((uint32_t)cb[p][i] << 16) + ((uint32_t)cb[p][i + L] << 17);
    }
    free(ca[p]);
    free(cb[p]);
  }
  return P[0].ctr1;
}

uint64_t orc_threshold(const uint16_t *const code_a[3], const uint16_t *const code_b[3],
                       const uint16_t *const mask_a[3], const uint16_t *const mask_b[3],
                       size_t L, size_t alloc_words, uint32_t threshold_a,
                       const uint32_t *seeds_own, const uint32_t *seeds_prev, uint64_t nonce,
                       uint64_t counter, uint64_t *const res_a[3], uint64_t *const res_b[3],
                       uint32_t *const lifted_mask_a_dbg[3]) {
  const size_t W = L / 64, BITS = 32;
  party_t P[3];
  init_parties(P, seeds_own, seeds_prev, nonce, counter);
  uint32_t *xa[3], *xb[3];
  uint16_t *ca[3], *cb[3];
  for (int p = 0; p < 3; p++) {
    xa[p] = malloc(L * 4); xb[p] = malloc(L * 4);
    ca[p] = malloc(2 * L * 2); cb[p] = malloc(2 * L * 2);
  }
  lift_mpc(P, mask_a, mask_b, L, alloc_words, xa, xb, ca, cb);
  /* lift_mul_sub (kernel.cu:181-195, 407-426) */
  for (int p = 0; p < 3; p++) {
    PAR_FOR(L)
    for (size_t i = 0; i < L; i++) {
      uint32_t ma = xa[p][i] - ((uint32_t)ca[p][i] << 16) - ((uint32_t)ca[p][i + L] << 17);
      uint32_t mb = xb[p][i] - ((uint32_t)cb[p][i] << 16) - ((uint32_t)cb[p][i + L] << 17);
      if (lifted_mask_a_dbg && lifted_mask_a_dbg[p]) lifted_mask_a_dbg[p][i] = ma;
      ma = ma * threshold_a - ((uint32_t)code_a[p][i] << 16);
      mb = mb * threshold_a - ((uint32_t)code_b[p][i] << 16);
      if (p == 0) ma -= 1;
      if (p == 1) mb -= 1;
      xa[p][i] = ma;
      xb[p][i] = mb;
    }
  }
  /* extract_msb */
  rs_t x1[3], x2[3], x3[3];
  for (int p = 0; p < 3; p++) {
    x1[p] = rs_new(BITS * W); x2[p] = rs_new(BITS * W); x3[p] = rs_new(BITS * W);
    transpose_u32(xa[p], W, 32, x1[p].a);
    transpose_u32(xb[p], W, 32, x1[p].b);
    split(p, x1[p].a, x1[p].b, x2[p].a, x2[p].b, x3[p].a, x3[p].b, BITS * W);
  }
  rs_t s[3], carry[3];
  for (int p = 0; p < 3; p++) {
    s[p] = rs_new((BITS - 1) * W);
    carry[p] = rs_new((BITS - 1) * W);
    xor_into(x2[p].a, x3[p].a, BITS * W);
    xor_into(x2[p].b, x3[p].b, BITS * W);
    for (size_t i = 0; i < (BITS - 1) * W; i++) {
      s[p].a[i] = x1[p].a[W + i] ^ x2[p].a[W + i];
      s[p].b[i] = x1[p].b[W + i] ^ x2[p].b[W + i];
    }
    xor_into(x1[p].a, x3[p].a, (BITS - 1) * W);
    xor_into(x1[p].b, x3[p].b, (BITS - 1) * W);
    and_pre(&P[p], x1[p].a, x1[p].b, x2[p].a, x2[p].b, carry[p].a, (BITS - 1) * W,
            CEIL_DIV((BITS - 1) * W, 8) * 8);
  }
  {
    uint64_t *A[3] = {carry[0].a, carry[1].a, carry[2].a}, *B[3] = {carry[0].b, carry[1].b, carry[2].b};
    send_recv_u64(P, A, B, (BITS - 1) * W);
  }
  for (int p = 0; p < 3; p++) {
    xor_into(carry[p].a, x3[p].a, (BITS - 1) * W);
    xor_into(carry[p].b, x3[p].b, (BITS - 1) * W);
  }
  /* c = plane 0 of x1 */
  for (int p = 0; p < 3; p++)
    and_pre(&P[p], s[p].a, s[p].b, carry[p].a, carry[p].b, x1[p].a, W, CEIL_DIV(W, 8) * 8);
  {
    uint64_t *A[3] = {x1[0].a, x1[1].a, x1[2].a}, *B[3] = {x1[0].b, x1[1].b, x1[2].b};
    send_recv_u64(P, A, B, W);
  }
  for (size_t k = 1; k < BITS - 2; k++) {
    uint64_t *sa[3], *sb[3];
    for (int p = 0; p < 3; p++) {
      uint64_t *ak_a = s[p].a + k * W, *ak_b = s[p].b + k * W;
      uint64_t *bk_a = carry[p].a + k * W, *bk_b = carry[p].b + k * W;
      xor_into(ak_a, x1[p].a, W);
      xor_into(ak_b, x1[p].b, W);
      xor_into(bk_a, x1[p].a, W);
      xor_into(bk_b, x1[p].b, W);
      and_pre(&P[p], ak_a, ak_b, bk_a, bk_b, s[p].a, W, CEIL_DIV(W, 8) * 8);
      sa[p] = s[p].a;
      sb[p] = s[p].b;
    }
    send_recv_u64(P, sa, sb, W);
    for (int p = 0; p < 3; p++) {
      xor_into(x1[p].a, s[p].a, W);
      xor_into(x1[p].b, s[p].b, W);
    }
  }
  for (int p = 0; p < 3; p++) {
    xor_into(x1[p].a, s[p].a + (BITS - 2) * W, W);
    xor_into(x1[p].b, s[p].b + (BITS - 2) * W, W);
    xor_into(x1[p].a, carry[p].a + (BITS - 2) * W, W);
    xor_into(x1[p].b, carry[p].b + (BITS - 2) * W, W);
    memcpy(res_a[p], x1[p].a, W * 8);
    memcpy(res_b[p], x1[p].b, W * 8);
  }
  uint64_t ctr = P[0].ctr1;
  int ok = 1;
  for (int p = 0; p < 3; p++) ok &= (P[p].ctr1 == ctr && P[p].ctr2 == ctr);
  for (int p = 0; p < 3; p++) {
    rs_free(&x1[p]); rs_free(&x2[p]); rs_free(&x3[p]); rs_free(&s[p]); rs_free(&carry[p]);
    free(xa[p]); free(xb[p]); free(ca[p]); free(cb[p]);
  }
  return ok ? ctr : (uint64_t)-1;
}
