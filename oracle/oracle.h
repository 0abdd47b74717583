/*
 * CPU oracle for the iris-mpc secret-shared matching path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product never links it.
 *
 * Every function restates the reference algorithm it follows, cited as
 * path:line relative to the worldcoin/iris-mpc repository root.
 *
 * Parity status per function family (how each is pinned):
 *   ChaCha block         : RFC 8439 s2.3.2 (20 rounds) + published 12-round
 *                          all-zero vector; reference has no KAT.
 *   SHA-256/HMAC/HKDF    : FIPS 180-2 "abc", RFC 5869 test cases 1 and 3.
 *   derive_seed          : equal to an independent RFC 5869 HKDF on Python's
 *                          hmac/hashlib built from the reference's literals
 *                          (KDF_SALT actor.rs:116; LE ikm words, 8-byte BE
 *                          nonce, LE output words, actor.rs:2914-2925); the
 *                          reference holds no derive_seed fixture.
 *   GR(2^16,4) arithmetic: Sage spec vectors (gr4 product/inverse), basis
 *                          identity <u_A,v_A> == coord0_B(S u_A * S v_A).
 *   Share encoding       : reconstruction invariant sum_i trick_dot == plaintext
 *                          dot (galois_engine.rs:713-782); per-share values vs
 *                          the reference generator: parity unpinned (external
 *                          ampc-secret-sharing RNG draw order).
 *   Iris bits/rotation   : example-data/all_rotations.txt, random_codes.txt.
 *   Limb GEMM+correction : brute-force u16 wrapping MAC (share_db.rs:1133-1198).
 *   3-party protocol     : opened value vs closed forms of tests/threshold.rs,
 *                          lift.rs, bitinject.rs; per-share intermediate values
 *                          vs the reference: parity unpinned (determined by the
 *                          ChaCha schedule restated here).
 */
#ifndef IRISMPC_ORACLE_H
#define IRISMPC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CODE_LEN 12800
#define ORC_MASK_LEN 6400
#define ORC_ROTATIONS 31
#define ORC_CODE_WORDS 200

/* ---------------- ChaCha (iris-mpc-gpu/src/rng/chacha.cu:74-183) ---------- */
void orc_chacha_block(const uint32_t in[16], uint32_t out[16], int rounds);
/* Keystream with the reference state layout: consts, key[8], 64-bit counter in
 * words 12-13, 64-bit nonce in words 14-15 (rng/chacha.rs:205-250). Writes
 * n_words u32 starting at block `counter`. */
/* Threads the keystream and the protocol simulator may use (default 1). */
void orc_set_threads(int n);
void orc_chacha_keystream(const uint32_t key[8], uint64_t nonce, uint64_t counter,
                          int rounds, uint32_t *out, size_t n_words);

/* ---------------- SHA-256 / HMAC / HKDF --------------------------------- */
void orc_sha256(const uint8_t *msg, size_t len, uint8_t out[32]);
void orc_hmac_sha256(const uint8_t *key, size_t klen, const uint8_t *msg, size_t mlen,
                     uint8_t out[32]);
void orc_hkdf_sha256(const uint8_t *salt, size_t slen, const uint8_t *ikm, size_t ilen,
                     const uint8_t *info, size_t infolen, uint8_t *okm, size_t okmlen);
/* server/actor.rs:2914-2925 with KDF_SALT (actor.rs:116). */
void orc_derive_seed(const uint32_t seed[8], uint64_t nonce, uint32_t out[8]);

/* ---------------- GR(2^16,4) = Z_2^16[x]/(x^4-x-1) ---------------------- */
void orc_gr_mul(const uint16_t a[4], const uint16_t b[4], uint16_t out[4]);
int orc_gr_inv(const uint16_t a[4], uint16_t out[4]);
void orc_gr_A_to_monomial(const uint16_t in[4], uint16_t out[4]);
void orc_gr_monomial_to_B(const uint16_t in[4], uint16_t out[4]);
void orc_gr_lagrange_at_zero(uint16_t out[12]); /* 3 elements for points 1, x, 1+x */
size_t orc_remap_new_to_old(size_t i);
size_t orc_remap_old_to_new(size_t i);
size_t orc_remap_new_to_mirrored(size_t i);  /* galois_engine.rs:276-289 */
/* galois_engine.rs:218-224,528-545: mirrored share (flip_imaginary for codes) */
void orc_share_mirrored(const uint16_t *in, size_t n, int flip_imaginary, uint16_t *out);

/* Encode plaintext (code, mask) into 3 additive-by-Lagrange GR shares.
 * rand supplies one random GR element (4 u16) per 4-coef block.
 * galois_engine.rs:319-364 (code), :366-397 (mask). */
void orc_encode_iris_code(const uint64_t code[200], const uint64_t mask[200],
                          const uint16_t *rand /* 3200*4 */, uint16_t *out /* 3*12800 */);
/* Trimmed mask: only the first 6400 coefs (b == 0 half) are produced
 * (galois_engine.rs:148-159 From<GaloisRingIrisCodeShare>). */
void orc_encode_mask_trimmed(const uint64_t mask[200], const uint16_t *rand /* 1600*4 */,
                             uint16_t *out /* 3*6400 */);
/* galois_engine.rs:23-40: multiply by Lagrange coefficient of party, to basis B. */
void orc_preprocess_query(int party_id, uint16_t *coefs, size_t n);
/* galois_engine.rs:42-52,479-488: 31 rotations, rotation 15 = identity. */
void orc_share_rotations(const uint16_t *coefs, size_t n, uint16_t *out /* 31*n */);
uint16_t orc_trick_dot(const uint16_t *a, const uint16_t *b, size_t n);

/* ---------------- plaintext iris (iris_db/iris.rs) ----------------------- */
int orc_iris_from_base64(const char *s, uint64_t out[200]);
void orc_iris_rotate_left(uint64_t code[200], int by);  /* by may be negative */
void orc_iris_all_rotations(const uint64_t code[200], uint64_t *out /* 31*200 */);
void orc_iris_distance(const uint64_t *c1, const uint64_t *m1, const uint64_t *c2,
                       const uint64_t *m2, uint32_t *hd, uint32_t *md);
/* iris.rs:587-601 */
void orc_iris_mirrored(const uint64_t code[200], const uint64_t mask[200], uint64_t out_code[200],
                       uint64_t out_mask[200]);
/* Integer predicate of the MPC circuit: msb_32(M*A - (c<<16) - 1) with
 * c = code dot (signed), M = mask dot (threshold_ring/cuda/kernel.cu:189-195,
 * 407-426; tests/threshold.rs:111-124). */
int orc_mpc_predicate(uint16_t code_dot, uint16_t mask_dot, uint32_t threshold_a);
/* Plaintext brute force of one query (all rotations) against n DB irises
 * ([n][200] codes and masks) at the integer predicate: out[d] = 0 / 1. */
void orc_plain_match(const uint64_t qc[200], const uint64_t qm[200], const uint64_t *dc,
                     const uint64_t *dm, size_t n, uint32_t threshold_a, uint8_t *out, int threads);
/* orc_mpc_predicate over n elements (out: 0 / 1). */
void orc_mpc_predicate_many(const uint16_t *code_dot, const uint16_t *mask_dot, size_t n,
                            uint32_t threshold_a, uint8_t *out);

/* ---------------- limbs, sums, share GEMM ------------------------------- */
/* share_db.rs:45-59 */
void orc_limbs(const uint16_t *v, size_t n, int8_t *l0, int8_t *l1);
uint32_t orc_limb_sum(const int8_t *l, size_t n);
/* Brute force u16 wrapping MAC: out[q*N + d] = sum_k q[q][k]*db[d][k] mod 2^16. */
void orc_u16_dot(const uint16_t *q, const uint16_t *db, size_t Qr, size_t N, size_t K,
                 uint16_t *out, int threads);
/* Limb-domain GEMM + dot/kernel.cu:16-28 correction, no masks (validates the
 * correction algebra against orc_u16_dot). */
void orc_limb_dot_corrected(const uint16_t *q, const uint16_t *db, size_t Qr, size_t N,
                            size_t K, uint16_t multiplier, uint16_t *out);
/* iris-mpc-cpu/src/protocol/ops.rs:238-347: one query (unrotated, preprocessed
 * code + trimmed mask shares) against n targets, all 31 rotations.
 * out[(t*31 + r)*2 + 0] = code dot, +1 = 2 * mask dot. OpenMP over targets. */
void orc_cpu_rotation_aware_dot(const uint16_t *q_code, const uint16_t *q_mask,
                                const uint16_t *db_code, const uint16_t *db_mask, size_t n,
                                uint16_t *out, int threads);

/* ---------------- ShareDB masking + reshare (share_db.rs:800-1036) ------ */
/* Per party p: out = (dot*mult) + F(s_own)[idx] - F(s_prev)[idx] (u16 views),
 * F = ChaCha12(seed, nonce, counter). Returns nothing; counter handled by
 * caller (masks use counter, OTP uses counter + L). */
void orc_mask_dot_share(uint16_t *dot, size_t n_elems, uint16_t multiplier,
                        const uint32_t seed_own[8], const uint32_t seed_prev[8],
                        uint64_t nonce, uint64_t counter);

/* ---------------- 3-party threshold protocol (threshold_ring/protocol.rs) */
/* Simulates all three parties of Circuits::compare_threshold_masked_many
 * (protocol.rs:2519-2546) bit-for-bit with the reference RNG counter schedule
 * and OTP'd exchanges. Inputs are replicated shares: party p holds
 * (a = x_p, b = x_{p-1}). L = elements (multiple of 64), alloc_words = the
 * Circuits buffer chunk size (bit-inject randomness is drawn for the whole
 * alloc). seeds_own[p] / seeds_prev[p] are the party's Circuits ChaCha seeds.
 * Outputs: result bit shares (plane 0 of lifted_shares_split1_result).
 * Returns the counter after the call (all six streams advance identically). */
uint64_t orc_threshold(const uint16_t *const code_a[3], const uint16_t *const code_b[3],
                       const uint16_t *const mask_a[3], const uint16_t *const mask_b[3],
                       size_t L, size_t alloc_words, uint32_t threshold_a,
                       const uint32_t *seeds_own /* 3*8 */, const uint32_t *seeds_prev /* 3*8 */,
                       uint64_t nonce, uint64_t counter,
                       uint64_t *const res_a[3], uint64_t *const res_b[3],
                       uint32_t *const lifted_mask_a_dbg[3] /* may be NULL */);

/* Lift only (protocol.rs:2041-2074 + finalize): returns per-party lifted u32
 * replicated shares of the mask (a, b) after correction. For lift.rs-style
 * tests. */
uint64_t orc_lift(const uint16_t *const mask_a[3], const uint16_t *const mask_b[3], size_t L,
                  size_t alloc_words, const uint32_t *seeds_own, const uint32_t *seeds_prev,
                  uint64_t nonce, uint64_t counter, uint32_t *const out_a[3],
                  uint32_t *const out_b[3]);

/* Open (actor.rs:2928-2997 + dot/kernel.cu:30-52,54-116): xor of three bit
 * shares, OR over rotations into bitmap[bit], bit = totalDb*(q/31) + db + offset
 * for q < query_len, db < real_chunk_len. */
void orc_open_bitmap(const uint64_t *r1, const uint64_t *r2, const uint64_t *r3,
                     size_t n_words, size_t chunk_len, size_t query_len, size_t offset,
                     size_t real_chunk_len, size_t total_db_len, uint64_t *bitmap);

/* ---------------- synthetic workload generator -------------------------- */
/* Deterministic ChaCha20 (key 0) generator, counter-addressed per iris so any
 * row can be regenerated independently (documented in the generator source). */
void orc_synth_iris(uint64_t dataset, uint64_t index, uint64_t code[200], uint64_t mask[200]);
void orc_synth_flip(const uint64_t code_in[200], uint64_t dataset, uint64_t index,
                    uint32_t n_flips, uint64_t code_out[200]);
void orc_synth_share_rand(uint64_t dataset, uint64_t index, uint16_t *code_rand /*3200*4*/,
                          uint16_t *mask_rand /*1600*4*/);
/* Full per-iris share generation: shares[p] code (12800) and mask (6400). */
void orc_synth_shares(const uint64_t code[200], const uint64_t mask[200], uint64_t dataset,
                      uint64_t index, uint16_t *code_sh /*3*12800*/, uint16_t *mask_sh /*3*6400*/);

#ifdef __cplusplus
}
#endif
#endif
