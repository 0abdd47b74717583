// Host computation of the GR(2^16,4) constants used by the query preprocessing
// kernel: Lagrange coefficients at zero and the monomial -> basis B matrix.
#include "galois.hpp"

#include <array>
#include <stdexcept>

namespace irismpc {
namespace {

using Mat = std::array<std::array<uint32_t, 4>, 4>;

// Change of basis A -> monomial (gr4_encoding_example.sage:4-11).
constexpr uint16_t kS[4][4] = {
    {1, 0, 0, 0}, {58082, 1, 0, 0}, {60579, 25194, 1, 0}, {17325, 51956, 57011, 1}};

// Inverse of an odd value mod 2^16 by Newton iteration.
uint32_t inv_odd(uint32_t v) {
  uint32_t x = v;
  for (int i = 0; i < 5; i++) x = (x * (2u - v * x)) & 0xffffu;
  return x;
}

// Solves M y = rhs mod 2^16 for a matrix whose determinant is odd.
std::array<uint16_t, 4> solve(Mat m, std::array<uint32_t, 4> rhs) {
  for (int col = 0; col < 4; col++) {
    int piv = col;
    while (piv < 4 && !(m[piv][col] & 1u)) piv++;
    if (piv == 4) throw std::runtime_error("GR(2^16,4): singular system");
    std::swap(m[piv], m[col]);
    std::swap(rhs[piv], rhs[col]);
    const uint32_t iv = inv_odd(m[col][col]);
    for (int j = 0; j < 4; j++) m[col][j] = (m[col][j] * iv) & 0xffffu;
    rhs[col] = (rhs[col] * iv) & 0xffffu;
    for (int r = 0; r < 4; r++) {
      if (r == col || !m[r][col]) continue;
      const uint32_t f = m[r][col];
      for (int j = 0; j < 4; j++) m[r][j] = (m[r][j] - f * m[col][j]) & 0xffffu;
      rhs[r] = (rhs[r] - f * rhs[col]) & 0xffffu;
    }
  }
  return {(uint16_t)rhs[0], (uint16_t)rhs[1], (uint16_t)rhs[2], (uint16_t)rhs[3]};
}

std::array<uint16_t, 4> gr_inv(const uint16_t a[4]) {
  Mat m{};
  for (int j = 0; j < 4; j++) {
    uint16_t e[4] = {0, 0, 0, 0}, col[4];
    e[j] = 1;
    gr_mul(a, e, col);
    for (int i = 0; i < 4; i++) m[i][j] = col[i];
  }
  return solve(m, {1, 0, 0, 0});
}

}  // namespace

void gr_mul(const uint16_t a[4], const uint16_t b[4], uint16_t out[4]) {
  const uint32_t a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
  const uint32_t b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
  // x^4 = x + 1, x^5 = x^2 + x, x^6 = x^3 + x^2
  const uint32_t d4 = a1 * b3 + a2 * b2 + a3 * b1, d5 = a2 * b3 + a3 * b2, d6 = a3 * b3;
  out[0] = (uint16_t)(a0 * b0 + d4);
  out[1] = (uint16_t)(a0 * b1 + a1 * b0 + d4 + d5);
  out[2] = (uint16_t)(a0 * b2 + a1 * b1 + a2 * b0 + d5 + d6);
  out[3] = (uint16_t)(a0 * b3 + a1 * b2 + a2 * b1 + a3 * b0 + d6);
}

QueryPrepConsts query_prep_consts(int party_id) {
  if (party_id < 0 || party_id > 2) throw std::runtime_error("party id must be 0..2");
  const uint16_t pts[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {1, 1, 0, 0}};
  QueryPrepConsts c{};
  // lambda_i = prod_{j != i} (0 - p_j) / (p_i - p_j)
  uint16_t lam[4] = {1, 0, 0, 0};
  for (int j = 0; j < 3; j++) {
    if (j == party_id) continue;
    uint16_t num[4], den[4], t[4];
    for (int k = 0; k < 4; k++) {
      num[k] = (uint16_t)(0u - pts[j][k]);
      den[k] = (uint16_t)(pts[party_id][k] - pts[j][k]);
    }
    gr_mul(lam, num, t);
    const auto di = gr_inv(den);
    gr_mul(t, di.data(), lam);
  }
  for (int k = 0; k < 4; k++) c.lagrange[k] = lam[k];
  // (S S^T)^-1, column by column
  Mat sst{};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      uint32_t s = 0;
      for (int k = 0; k < 4; k++) s += (uint32_t)kS[i][k] * kS[j][k];
      sst[i][j] = s & 0xffffu;
    }
  for (int j = 0; j < 4; j++) {
    std::array<uint32_t, 4> e{0, 0, 0, 0};
    e[j] = 1;
    const auto col = solve(sst, e);
    for (int i = 0; i < 4; i++) c.to_b[4 * i + j] = col[i];
  }
  return c;
}

}  // namespace irismpc
