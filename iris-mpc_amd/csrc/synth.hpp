// Synthetic iris workloads: plaintext codes and masks, planted bit flips and
// the three parties' GR(2^16,4) shares, generated from ChaCha20 streams under
// the all-zero key (the data the benchmarks and full-size tests run on).
#pragma once
#include "common.hpp"

namespace irismpc {

constexpr int kIrisWords = 200;  // u64 words of a 12,800-bit code or mask

// Host: plaintext iris (dataset, index) and n_flips distinct flipped code bits.
void synth_iris_host(uint64_t dataset, uint64_t index, uint64_t code[kIrisWords],
                     uint64_t mask[kIrisWords]);
void synth_flip_host(uint64_t code[kIrisWords], uint64_t dataset, uint64_t index, uint32_t n_flips);

// Device: plaintext irises index0 .. index0 + n - 1 of a dataset ([n][200] u64).
void launch_synth_plain(uint64_t dataset, uint64_t index0, size_t n, uint64_t* codes,
                        uint64_t* masks, hipStream_t s);
// Device: shares of n plaintext irises for all three parties; iris i draws its
// sharing randomness from (rand_dataset, rand_index0 + i). code_out[p] is
// [n][12800] u16, mask_out[p] is [n][6400] u16 (trimmed mask).
void launch_synth_shares(const uint64_t* codes, const uint64_t* masks, size_t n,
                         uint64_t rand_dataset, uint64_t rand_index0, uint16_t* const code_out[3],
                         uint16_t* const mask_out[3], hipStream_t s);

}  // namespace irismpc
