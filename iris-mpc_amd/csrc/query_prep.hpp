// Query preprocessing (Lagrange factor, basis B, 31 rotations) on the GPU.
#pragma once
#include "common.hpp"

namespace irismpc {
// src: batch x K raw shares of party party_id (device); dst: rows_out x K.
// Rows of queries beyond `batch` (padding up to rows_out) are zero.
// mirror: 0 as given; 1 mirrored mask shares; 2 mirrored code shares
// (imaginary half negated).
void launch_query_prep(const uint16_t* src, int batch, int K, int party_id, uint16_t* dst,
                       int rows_out, hipStream_t s, int mirror = 0);
}  // namespace irismpc
