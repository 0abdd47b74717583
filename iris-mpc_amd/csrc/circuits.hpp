// Threshold-comparison circuits on replicated shares (3 parties, one GPU).
#pragma once
#include "anon.hpp"
#include "common.hpp"

namespace irismpc {

// One party's Circuits state (device pointers). W = words per bit plane
// (elements / 64). The kernels run the three parties together (one wave per
// party in a block), and a message to the next party reaches its wave through
// LDS: that is the exchange which replaces the reference's NCCL send/recv.
struct CircuitBufs {
  const uint16_t* code_a;
  const uint16_t* code_b;
  const uint16_t* mask_a;
  const uint16_t* mask_b;
  uint64_t *c_a, *c_b;      // 2 W: lift carries (bit 16, bit 17)
  uint16_t* corr_a;  // 128 W: injected carries, a share (the b share is the previous party's a)
  uint64_t *res_a, *res_b;  // W: result bit shares (plane 0)
  uint32_t *lifted_a, *lifted_b;  // optional (debug): corrected lifted mask
  ChaChaKey key_own, key_prev;    // Circuits ChaCha seeds (chacha1 / chacha2)
};

struct Trio {
  CircuitBufs p[3];
  size_t W;
  uint64_t nonce;
};
// The trio's seeds are pairwise: party p's key_prev is party p - 1's key_own.
// Each pairwise ChaCha12 block is computed once, by the wave of the party that
// owns its key, and read by the next party's wave from LDS.

// Block counters of every RNG draw of compare_threshold_masked_many, following
// the reference schedule (protocol.rs:2519-2546): all six ChaCha streams of the
// three parties advance identically, OTP draws included.
struct CircuitSchedule {
  uint64_t lift_csa;            // packed AND over 16 W words
  uint64_t lift_first;          // first full adder
  uint64_t lift_round[15];      // ripple rounds k = 1..14 (index k)
  uint64_t lift_final;          // carry-out AND
  uint64_t inject;              // rand_01 / rand_02 / rand_12 base
  uint64_t msb_csa;             // packed AND over 31 W words
  uint64_t msb_first;
  uint64_t msb_round[30];       // k = 1..29
  uint64_t end;                 // counter after the call
  static CircuitSchedule make(uint64_t start, size_t W, size_t alloc_words);
};

// ChaCha12 block counters of one adder: the carry-save AND of plane j uses
// csa + j W / 8 (+ word / 8), ripple round k uses c[k] (lift: c[15] is the
// carry-out AND).
struct RippleCtrs {
  uint64_t csa;
  uint64_t c[32];
};

// One launch per adder, all three parties: msb = false is
// the lift (mask shares -> carries in c planes 0 and 1), msb = true the
// lift_mul_sub and the MSB adder (-> res). The bit planes stay in registers.
void launch_adder_word(const Trio& t, bool msb, uint32_t threshold_a, const RippleCtrs& ctrs,
                       hipStream_t s);
// True when every party's b views (code and mask) are the previous party's a
// buffers, as the zero-copy reshare hands them over (and the optional lift
// capture is set for all or none).
bool trio_b_is_prev(const Trio& t);
// lift, bit injection, lift_mul_sub and MSB adder in one launch (needs
// trio_b_is_prev): carries and corrections stay in registers and LDS; the
// result shares equal those of the three separate launches.
void launch_compare_fused(const Trio& t, uint32_t threshold_a, const RippleCtrs& lift, uint64_t inject_ctr,
                          const RippleCtrs& msb, hipStream_t s);
// Bit injection of the two lift carries (all three parties): the a share
// corr_a of every element's corrections (the b share is the previous party's
// a share), the message steps inside one launch.
void launch_inject(const Trio& t, uint64_t ctr, hipStream_t s);

// Open + bitmap (actor.rs:2928-2997, dot/kernel.cu:30-116). Party views:
// a = own result.a, b = own result.b, c = prev party's result.b (received).
struct OpenArgs {
  const uint64_t* a;
  const uint64_t* b;
  const uint64_t* c;
  uint64_t* bitmap;
  size_t n_words;        // W
  size_t chunk_len;      // padded chunk width (multiple of 64)
  size_t query_len;      // real batch * 31
  size_t offset;         // global DB offset of the chunk
  size_t real_chunk_len;
  size_t total_db_len;
  const uint32_t* slot_of;  // optional: bitmap row of batch query j
  AnonCapture anon;         // optional match-distance capture
  DotViews dots;
};
void launch_open(const OpenArgs& a, hipStream_t s);

}  // namespace irismpc
