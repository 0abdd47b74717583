/* Oracle: Galois ring GR(2^16, 4) = Z_2^16[x]/(x^4 - x - 1) and the 3-party
 * Shamir share encoding of iris codes. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Arithmetic restates iris-mpc-common/src/sage/symmetric-gr-shamir_sage.sage
 * (prod_monom :150-165, legrange_coeffs :388-405, base points 1, x, 1+x
 * :373-377) and the change-of-basis matrices of
 * iris-mpc-common/src/sage/gr4_encoding_example.sage:4-11,77-78
 * (A -> monomial = S, monomial -> B = (S S^T)^-1).
 * Encoding restates iris-mpc-common/src/galois_engine.rs:319-397
 * (encode_iris_code / encode_mask_code), :23-40 (preprocess_coefs),
 * :256-273 (index remaps), :42-52 (rotations), :55-61 (trick_dot).
 * The external ampc-secret-sharing crate (rev cec37f5b) that implements
 * encode_3_mat is not vendored; the degree-1 sharing f(X) = s + r X evaluated
 * at the base points is restated from the Sage spec. */
#include <string.h>
#include "../oracle.h"

static const uint16_t S16[4][4] = {
    {1, 0, 0, 0}, {58082, 1, 0, 0}, {60579, 25194, 1, 0}, {17325, 51956, 57011, 1}};

void orc_gr_mul(const uint16_t a[4], const uint16_t b[4], uint16_t out[4]) {
  uint32_t a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
  uint32_t b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
  uint16_t c0 = (uint16_t)(a3 * b1 + a2 * b2 + a1 * b3 + a0 * b0);
  uint16_t c1 = (uint16_t)(a3 * b2 + a2 * b3 + a3 * b1 + a2 * b2 + a1 * b3 + a1 * b0 + a0 * b1);
  uint16_t c2 = (uint16_t)(a3 * b3 + a3 * b2 + a2 * b3 + a2 * b0 + a1 * b1 + a0 * b2);
  uint16_t c3 = (uint16_t)(a3 * b3 + a3 * b0 + a2 * b1 + a1 * b2 + a0 * b3);
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Solve M x = y mod 2^16 for a 4x4 M with odd determinant (odd pivots). */
static int solve4(uint16_t Min[4][4], const uint16_t y[4], uint16_t x[4]) {
  uint32_t M[4][5];
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) M[i][j] = Min[i][j];
    M[i][4] = y[i];
  }
  for (int col = 0; col < 4; col++) {
    int piv = -1;
    for (int r = col; r < 4; r++)
      if (M[r][col] & 1) { piv = r; break; }
    if (piv < 0) return -1;
    if (piv != col)
      for (int j = 0; j < 5; j++) { uint32_t t = M[col][j]; M[col][j] = M[piv][j]; M[piv][j] = t; }
    /* inverse of odd pivot mod 2^16 via Newton */
    uint32_t p = M[col][col] & 0xffff, inv = p;
    for (int it = 0; it < 5; it++) inv = (inv * (2u - p * inv)) & 0xffff;
    for (int j = 0; j < 5; j++) M[col][j] = (M[col][j] * inv) & 0xffff;
    for (int r = 0; r < 4; r++) {
      if (r == col) continue;
      uint32_t f = M[r][col];
      for (int j = 0; j < 5; j++) M[r][j] = (M[r][j] - f * M[col][j]) & 0xffff;
    }
  }
  for (int i = 0; i < 4; i++) x[i] = (uint16_t)M[i][4];
  return 0;
}

int orc_gr_inv(const uint16_t a[4], uint16_t out[4]) {
  /* column j of the multiplication matrix = a * x^j */
  uint16_t M[4][4];
  for (int j = 0; j < 4; j++) {
    uint16_t e[4] = {0, 0, 0, 0}, col[4];
    e[j] = 1;
    orc_gr_mul(a, e, col);
    for (int i = 0; i < 4; i++) M[i][j] = col[i];
  }
  const uint16_t one[4] = {1, 0, 0, 0};
  return solve4(M, one, out);
}

void orc_gr_A_to_monomial(const uint16_t in[4], uint16_t out[4]) {
  uint16_t r[4];
  for (int i = 0; i < 4; i++) {
    uint32_t s = 0;
    for (int j = 0; j < 4; j++) s += (uint32_t)S16[i][j] * in[j];
    r[i] = (uint16_t)s;
  }
  memcpy(out, r, 8);
}

static int g_B_ready = 0;
static uint16_t g_monom_to_B[4][4];

static void init_B(void) {
  if (g_B_ready) return;
  uint16_t SST[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      uint32_t s = 0;
      for (int k = 0; k < 4; k++) s += (uint32_t)S16[i][k] * S16[j][k];
      SST[i][j] = (uint16_t)s;
    }
  /* invert column by column */
  uint16_t inv[4][4];
  for (int j = 0; j < 4; j++) {
    uint16_t e[4] = {0, 0, 0, 0}, x[4];
    e[j] = 1;
    uint16_t Mc[4][4];
    memcpy(Mc, SST, sizeof(Mc));
    solve4(Mc, e, x);
    for (int i = 0; i < 4; i++) inv[i][j] = x[i];
  }
  memcpy(g_monom_to_B, inv, sizeof(inv));
  g_B_ready = 1;
}

void orc_gr_monomial_to_B(const uint16_t in[4], uint16_t out[4]) {
  init_B();
  uint16_t r[4];
  for (int i = 0; i < 4; i++) {
    uint32_t s = 0;
    for (int j = 0; j < 4; j++) s += (uint32_t)g_monom_to_B[i][j] * in[j];
    r[i] = (uint16_t)s;
  }
  memcpy(out, r, 8);
}

/* Base points of the 3 parties (ids 1..3): 1, x, 1+x. */
static const uint16_t BASE_PTS[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {1, 1, 0, 0}};

void orc_gr_lagrange_at_zero(uint16_t out[12]) {
  for (int i = 0; i < 3; i++) {
    uint16_t coeff[4] = {1, 0, 0, 0};
    for (int j = 0; j < 3; j++) {
      if (j == i) continue;
      uint16_t num[4], den[4], dinv[4], t[4];
      for (int k = 0; k < 4; k++) {
        num[k] = (uint16_t)(0 - BASE_PTS[j][k]);
        den[k] = (uint16_t)(BASE_PTS[i][k] - BASE_PTS[j][k]);
      }
      orc_gr_mul(coeff, num, t);
      orc_gr_inv(den, dinv);
      orc_gr_mul(t, dinv, coeff);
    }
    memcpy(out + 4 * i, coeff, 8);
  }
}

size_t orc_remap_new_to_old(size_t i) {
  size_t b = i / 6400, w = i % 6400 / 3200, r1 = i % 3200 / 800, c = i % 800 / 4, r2 = i % 4;
  size_t r = r2 * 4 + r1;
  return 800 * r + c * 4 + w * 2 + b;
}

size_t orc_remap_old_to_new(size_t i) {
  size_t r = i / 800, c = i % 800 / 4, w = i % 4 / 2, b = i % 2;
  size_t r1 = r % 4, r2 = r / 4;
  return b * 6400 + w * 3200 + r1 * 800 + c * 4 + r2;
}

/* galois_engine.rs:276-289: column c of each 800-coefficient row is mirrored
 * inside its half (0..99 and 100..199), other index parts kept. */
size_t orc_remap_new_to_mirrored(size_t i) {
  size_t b = i / 6400, w = i % 6400 / 3200, r1 = i % 3200 / 800, c = i % 800 / 4, r2 = i % 4;
  size_t fc = c < 100 ? 99 - c : 199 - (c - 100);
  return b * 6400 + w * 3200 + r1 * 800 + fc * 4 + r2;
}

/* GaloisRingIrisCodeShare::mirrored (galois_engine.rs:528-545) and the trimmed
 * mask's mirrored (:218-224): out[mirror(i)] = in[i]; with flip_imaginary the
 * coefficients of the imaginary half (i >= 6400) are first negated. The
 * reference negates by subtracting from a fixed sharing of zero
 * (flipped_imaginary_coeffs, :291-300, randomness StdRng(0)); any sharing of
 * zero gives a sharing of the same negated secret, and this restatement uses
 * the all-zero one: the opened dot products are the same. */
void orc_share_mirrored(const uint16_t *in, size_t n, int flip_imaginary, uint16_t *out) {
  for (size_t i = 0; i < n; i++) {
    uint16_t v = in[i];
    if (flip_imaginary && i >= ORC_CODE_LEN / 2) v = (uint16_t)(0u - v);
    out[orc_remap_new_to_mirrored(i)] = v;
  }
}

static inline int get_bit(const uint64_t *a, size_t i) { return (int)((a[i / 64] >> (i % 64)) & 1); }

/* Degree-1 sharing of a monomial-basis element s with randomness r at the
 * three base points: share_j = s + r * p_j. */
static void share3(const uint16_t s[4], const uint16_t r[4], uint16_t sh[3][4]) {
  for (int j = 0; j < 3; j++) {
    uint16_t rp[4];
    orc_gr_mul(r, BASE_PTS[j], rp);
    for (int k = 0; k < 4; k++) sh[j][k] = (uint16_t)(s[k] + rp[k]);
  }
}

void orc_encode_iris_code(const uint64_t code[200], const uint64_t mask[200],
                          const uint16_t *rand, uint16_t *out) {
  for (size_t i = 0; i < ORC_CODE_LEN; i += 4) {
    uint16_t el[4];
    for (int k = 0; k < 4; k++) {
      size_t o = orc_remap_new_to_old(i + k);
      uint16_t m = (uint16_t)get_bit(mask, o), c = (uint16_t)get_bit(code, o);
      el[k] = (uint16_t)(m - 2 * (c & m));
    }
    uint16_t mono[4], sh[3][4];
    orc_gr_A_to_monomial(el, mono);
    share3(mono, rand + i, sh);
    for (int j = 0; j < 3; j++) memcpy(out + (size_t)j * ORC_CODE_LEN + i, sh[j], 8);
  }
}

void orc_encode_mask_trimmed(const uint64_t mask[200], const uint16_t *rand, uint16_t *out) {
  for (size_t i = 0; i < ORC_MASK_LEN; i += 4) {
    uint16_t el[4];
    for (int k = 0; k < 4; k++) el[k] = (uint16_t)get_bit(mask, orc_remap_new_to_old(i + k));
    uint16_t mono[4], sh[3][4];
    orc_gr_A_to_monomial(el, mono);
    share3(mono, rand + i, sh);
    for (int j = 0; j < 3; j++) memcpy(out + (size_t)j * ORC_MASK_LEN + i, sh[j], 8);
  }
}

void orc_preprocess_query(int party_id, uint16_t *coefs, size_t n) {
  uint16_t lag[12];
  orc_gr_lagrange_at_zero(lag);
  for (size_t i = 0; i < n; i += 4) {
    uint16_t t[4];
    orc_gr_mul(coefs + i, lag + 4 * party_id, t);
    orc_gr_monomial_to_B(t, coefs + i);
  }
}

void orc_share_rotations(const uint16_t *coefs, size_t n, uint16_t *out) {
  const size_t CH = 800;
  for (int t = 0; t < ORC_ROTATIONS; t++) {
    long shift = 4L * (15 - t); /* net rotate_left by (15 - t) columns */
    uint16_t *o = out + (size_t)t * n;
    for (size_t base = 0; base < n; base += CH)
      for (size_t p = 0; p < CH; p++) {
        long src = ((long)p + shift) % (long)CH;
        if (src < 0) src += CH;
        o[base + p] = coefs[base + (size_t)src];
      }
  }
}

uint16_t orc_trick_dot(const uint16_t *a, const uint16_t *b, size_t n) {
  uint16_t s = 0;
  for (size_t i = 0; i < n; i++) s = (uint16_t)(s + (uint32_t)a[i] * b[i]);
  return s;
}
