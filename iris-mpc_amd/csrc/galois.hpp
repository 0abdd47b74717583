// GR(2^16, 4) = Z_2^16[x]/(x^4 - x - 1) constants for query preprocessing.
//
// preprocess_coefs (iris-mpc-common/src/galois_engine.rs:23-40) multiplies every
// 4-coefficient element of a party's query share by that party's Lagrange
// coefficient at zero for the base points 1, x, 1+x and converts the product
// to basis B = (S S^T)^-1 (iris-mpc-common/src/sage/gr4_encoding_example.sage
// :4-11,77-78, symmetric-gr-shamir_sage.sage:363-406).
#pragma once
#include <cstdint>

namespace irismpc {

struct QueryPrepConsts {
  uint16_t lagrange[4];   // Lagrange coefficient of this party (monomial basis)
  uint16_t to_b[16];      // row-major 4x4 monomial -> basis B matrix mod 2^16
};

QueryPrepConsts query_prep_consts(int party_id);

// Monomial-basis product (Sage prod_monom).
void gr_mul(const uint16_t a[4], const uint16_t b[4], uint16_t out[4]);

}  // namespace irismpc
