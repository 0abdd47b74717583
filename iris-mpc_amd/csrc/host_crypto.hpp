// Host-side seed derivation: HKDF-SHA256 as used by derive_seed
// (iris-mpc-gpu/src/server/actor.rs:2914-2925, KDF_SALT at actor.rs:116).
#pragma once
#include <cstddef>
#include <cstdint>

namespace irismpc {
void sha256(const uint8_t* msg, size_t len, uint8_t out[32]);
void hmac_sha256(const uint8_t* key, size_t klen, const uint8_t* msg, size_t mlen, uint8_t out[32]);
// HKDF-Extract then a single Expand block (okm <= 32 bytes suffices here).
void hkdf_sha256_32(const uint8_t* salt, size_t slen, const uint8_t* ikm, size_t ilen,
                    const uint8_t* info, size_t infolen, uint8_t out[32]);
void derive_seed(const uint32_t seed[8], uint64_t nonce, uint32_t out[8]);
}  // namespace irismpc
