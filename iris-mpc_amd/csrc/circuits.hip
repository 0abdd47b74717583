// Bit-sliced replicated-share circuits for gfx950.
//
// Restates threshold_ring/protocol.rs (lift_mpc :2041-2074,
// binary_add_3_get_two_carries :2137-2241, bit_inject :1302-1643,
// lift_mul_sub :1831-1872, extract_msb/binary_add_3_get_msb :2243-2359) and
// threshold_ring/cuda/kernel.cu. Design for one GPU:
//  * a block holds the same elements of all three parties, one wave per
//    party; a message to the next party goes to that party's wave through
//    LDS, and a barrier is the exchange;
//  * the 16 / 32 bit planes of the adders are formed by in-register bit
//    transposes and never leave registers;
//  * x1/x2/x3 of `split` are never materialised: for party p they are fixed
//    selections of the (a, b) planes, so x2^x3, x1^x3 and the carry-save AND
//    collapse to one expression per party (bit-identical);
//  * correlated randomness is generated in registers from the block counter
//    the reference would have used (no keystream buffers).
#include "circuits.hpp"

#include "chacha.hpp"

namespace irismpc {

CircuitSchedule CircuitSchedule::make(uint64_t c, size_t W, size_t alloc) {
  CircuitSchedule s{};
  const uint64_t b1 = ceil_div(W, 8);
  s.lift_csa = c;
  c += 2 * ceil_div(16 * W, 8);  // AND draw + OTP of the 16-plane exchange
  s.lift_first = c;
  c += 2 * b1;
  for (int k = 1; k <= 14; k++) {
    s.lift_round[k] = c;
    c += 2 * b1;
  }
  s.lift_final = c;
  c += 2 * b1;
  s.inject = c;
  c += 4 * alloc + 4 * W;  // rand buffers cover the whole alloc; u16 OTP of 128 W
  s.msb_csa = c;
  c += 2 * ceil_div(31 * W, 8);
  s.msb_first = c;
  c += 2 * b1;
  for (int k = 1; k <= 29; k++) {
    s.msb_round[k] = c;
    c += 2 * b1;
  }
  s.end = c;
  return s;
}

// ------------------------------------------------------------ carry-save AND
// and_pre(x1^x3, x2^x3) for the split of party p (kernel.cu:39-42,205-235):
//   p0: x1=(a,0) x2=0 x3=(0,b)  -> a & b
//   p1: x1=(0,b) x2=(a,0) x3=0  -> a & b
//   p2: x1=0 x2=(0,b) x3=(a,0)  -> a ^ (a & b)
__device__ __forceinline__ uint32_t csa_and(int p, uint32_t a, uint32_t b) {
  const uint32_t n = p == 2 ? ~0u : 0u;  // p is wave-uniform: a mask, not a select
  return a & (b ^ n);
}

// All ones for party q, zero otherwise. p is wave-uniform, so this is one
// scalar load from a constant table; a mask the compiler cannot see to be all
// or nothing stays an AND that folds into a three-input bit op (computed as
// p == q ? ~0 : 0 it becomes a per-lane select plus the xor).
__constant__ uint32_t kPartyEq[3] = {~0u, 0u, 0u};
__device__ __forceinline__ uint32_t party_mask(int p, int q) { return kPartyEq[(p - q + 3) % 3]; }

// ------------------------------------------------------------ ripple helpers
// Local share of the AND of replicated shares (x_a, x_b) and (y_a, y_b) plus
// the zero-share r (protocol.rs:610-710). In ripple round k, a_k is plane
// k + 1 and b_k the carry-save plane k ^ x3 plane k (x3 = (0, b) for p0,
// (a, 0) for p2, 0 for p1).
__device__ __forceinline__ uint32_t and_pre(uint32_t xa, uint32_t xb, uint32_t ya, uint32_t yb,
                                            uint32_t r) {
  return (xa & ya) ^ (xb & ya) ^ (xa & yb) ^ r;
}

// ------------------------------------------------------------ bit injection
// 128 W elements e: bits of the two lift carries (e < 64 W: bit 16,
// e >= 64 W: bit 17); ChaCha12 block g of each stream covers elements 32 g .. 32 g + 31.
// kernel.cu:526-629 with the message flow P1 -> P2 -> P0 and P0 -> P1.

// Block per 2048 elements, one wave per party: lane l draws block 64 g + l of
// its party's own stream once (block g covers elements 32 g .. 32 g + 31 as
// the u16 view of its keystream); the party's other stream (their_rng) is the
// previous party's own stream at the same counter, read from LDS. The
// keystreams go through LDS in element order, then each lane handles 8
// consecutive elements per pass. P0 and P1 write their a shares, P1 passes
// its message x to P2's wave through LDS, and after one barrier P2 writes its
// a share; every b share is the previous party's a share and is not stored.
__device__ __forceinline__ void u16x8_load(const uint16_t* p, uint16_t v[8]) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = (uint16_t)(w[i >> 1] >> (16 * (i & 1)));
}

__device__ __forceinline__ void u16x8_store(uint16_t* p, const uint16_t v[8]) {
  uint32_t w[4];
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = (uint32_t)v[2 * i] | ((uint32_t)v[2 * i + 1] << 16);
  *reinterpret_cast<uint4*>(p) = uint4{w[0], w[1], w[2], w[3]};
}

__global__ __launch_bounds__(192) void inject_kernel(const Trio t, uint64_t ctr) {
  // one party per wave: p is wave-uniform, so its keys and pointers stay scalar
  const int p = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63), prev = (p + 2) % 3;
  const CircuitBufs& P = t.p[p];
  // [party][own stream: u16 pairs in element order, 16-byte granules swizzled]:
  // granule j of lane w's block sits at granule 4 w + (j ^ (w & 3)), so the
  // lanes' 16-byte writes of one granule spread over more banks
  __shared__ uint32_t ks[3][1024];
  __shared__ uint32_t msg[1024];  // P1's message x to P2, u16 pairs in element order
  const size_t n = 128 * t.W;
  const size_t e_base = (size_t)blockIdx.x * 2048;
  {
    uint32_t ko[16];
    chacha_block<12>(P.key_own, t.nonce, ctr + (uint64_t)blockIdx.x * 64 + lane, ko);  // my_rng (chacha1)
#pragma unroll
    for (int i = 0; i < 16; i += 4)
      *reinterpret_cast<uint4*>(&ks[p][lane * 16 + 4 * ((i / 4) ^ (lane & 3))]) =
          uint4{ko[i], ko[i + 1], ko[i + 2], ko[i + 3]};
  }
  __syncthreads();
  const uint16_t* own16 = reinterpret_cast<const uint16_t*>(ks[p]);
  const uint16_t* prv16 = reinterpret_cast<const uint16_t*>(ks[prev]);  // their_rng (chacha2)
  // u16 offset of the 8 elements starting at element el (el a multiple of 8:
  // one granule) under the swizzle above
  auto kso = [](int el) {
    const int g = el >> 3, w = g >> 2;
    return 8 * (4 * w + ((g & 3) ^ (w & 3)));
  };
  // Party p's b share of a correction is party p - 1's a share (replicated
  // sharing: the reference's P1 writes its b from P0's message y, P2 its b
  // from P1's x, P0 its b from P2's z, each equal to the sender's a), so only
  // the a shares are written; consumers read b from the previous party's a.
  // Of the three messages only P1 -> P2 (x) feeds an a share.
  uint16_t* to_p2 = reinterpret_cast<uint16_t*>(msg);
  const uint16_t* inbox = reinterpret_cast<const uint16_t*>(msg);
  // step 1: P0 and P1 write corr_a; P1 sends x to P2
  if (p != 2) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int el = 8 * (k * 64 + lane);
      const size_t e = e_base + el;
      if (e >= n) break;
      uint16_t own[8], prv[8], o1[8], o2[8];
      u16x8_load(own16 + kso(el), own);
      u16x8_load(prv16 + kso(el), prv);
      if (p == 0) {
        // rand_01 = my rng, rand_02 = their rng; in_b = party 0's b share
        const uint32_t bits = (uint32_t)(P.c_b[e >> 6] >> (e & 63)) & 0xffu;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint16_t b = (uint16_t)((bits >> i) & 1u), r01 = own[i], r02 = prv[i];
          const uint16_t y = (uint16_t)(b * r01 - r02);
          o2[i] = (uint16_t)(r01 - (y << 1));
        }
      } else {
        // rand_01 = their rng, rand_12 = my rng
        const uint32_t bits = (uint32_t)((P.c_a[e >> 6] ^ P.c_b[e >> 6]) >> (e & 63)) & 0xffu;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint16_t r01 = prv[i], r12 = own[i];
          const uint16_t x = (uint16_t)(((bits >> i) & 1u) - r01);
          o1[i] = x;
          o2[i] = (uint16_t)(x - (r12 << 1));
        }
        u16x8_store(to_p2 + el, o1);
      }
      u16x8_store(P.corr_a + e, o2);
    }
  }
  __syncthreads();
  // step 2: P2 writes corr_a from P1's x
  if (p == 2) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int el = 8 * (k * 64 + lane);
      const size_t e = e_base + el;
      if (e >= n) break;
      uint16_t own[8], prv[8], in[8], o2[8];
      u16x8_load(own16 + kso(el), own);
      u16x8_load(prv16 + kso(el), prv);
      u16x8_load(inbox + el, in);
      // rand_12 = their rng, rand_02 = my rng; in_a = party 2's a share
      const uint32_t bits = (uint32_t)(P.c_a[e >> 6] >> (e & 63)) & 0xffu;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const uint16_t r12 = prv[i], r02 = own[i], x = in[i], b = (uint16_t)((bits >> i) & 1u);
        const uint16_t z = (uint16_t)(b * x - r12);
        o2[i] = (uint16_t)(-((r02 + z) << 1) + b);
      }
      u16x8_store(P.corr_a + e, o2);
    }
  }
}

// ------------------------------------------------------------ word adders
// One launch per adder, all three parties: the lift (16 planes of the masked
// mask share, carry-out into c planes 0 and 1) and the MSB stage (the 32
// planes of m = (mask - c1<<16 - c2<<17) * A - code<<16 - [party], result into
// res). A block holds the same 32 words of the three parties, one wave per
// party and one lane per half word (32 elements): the lane loads its
// elements, forms the values and turns them into 32-bit plane slices in
// registers (one 32 x 32 bit transpose per share), so no plane goes through
// HBM. Then step 0 is the carry-save AND of plane 0 and step s the
// carry-save AND of plane s with ripple round s - 1 (protocol.rs:2137-2359,
// kernel.cu:205-235); each step's AND outputs reach the next party through
// LDS.
//
// The ChaCha12 block pair behind one AND of 8 words (F(own) ^ F(prev),
// counter as in the schedule) is computed once per 16-lane group: for a batch
// of 8 steps (16 gates) lane i of a group computes gate i and LDS hands every
// lane its half word of all 16. A wave computes F(own) only: its F(prev) is
// the previous party's F(own) (pairwise seeds), which that party's wave has
// put in LDS, so the trio evaluates 3 distinct blocks per gate, not 6. The
// lift's unused slot (step 0 has no ripple round) carries the carry-out AND.

// Transpose a 32 x 32 bit matrix in place: afterwards bit i of x[j] is bit j
// of the old x[i] (recursive block swap, 5 stages).
// Stage S swaps, between x[k] and x[k + S], the S-bit fields that the mask M
// marks in x[k + S] with their neighbours above them in x[k]. The 16- and 8-bit
// stages move whole bytes: one v_perm_b32 per output, whose selector bytes
// 0-3 pick the bytes of the second operand and 4-7 those of the first. The
// others are a shift and a bit-field insert per output.
template <int S, uint32_t M, int N = 32>
__device__ __forceinline__ void bit_transpose_stage(uint32_t x[]) {
#pragma unroll
  for (int k = 0; k < N; k++)
    if ((k & S) == 0) {
      const uint32_t a = x[k], b = x[k + S];
      if constexpr (S == 16) {
        x[k] = __builtin_amdgcn_perm(b, a, 0x05040100u);      // a.lo | b.lo << 16
        x[k + S] = __builtin_amdgcn_perm(b, a, 0x07060302u);  // a.hi | b.hi << 16
      } else if constexpr (S == 8) {
        x[k] = __builtin_amdgcn_perm(b, a, 0x06020400u);      // bytes a0 b0 a2 b2
        x[k + S] = __builtin_amdgcn_perm(b, a, 0x07030501u);  // bytes a1 b1 a3 b3
      } else {
        static_assert(M << S == ~M, "the fields and their neighbours tile the word");
        x[k] = (a & M) | ((b << S) & ~M);
        x[k + S] = ((a >> S) & M) | (b & ~M);
      }
    }
}

__device__ __forceinline__ void bit_transpose32(uint32_t x[32]) {
  bit_transpose_stage<16, 0x0000FFFFu>(x);
  bit_transpose_stage<8, 0x00FF00FFu>(x);
  bit_transpose_stage<4, 0x0F0F0F0Fu>(x);
  bit_transpose_stage<2, 0x33333333u>(x);
  bit_transpose_stage<1, 0x55555555u>(x);
}

// The same for 16-bit values (x[i] < 2^16): afterwards bit i of x[j], j < 16,
// is bit j of the old x[i] (x[16 ..] are left undefined). Values i and i + 16
// share a register from the start, so the first stage of bit_transpose32 is
// the packing itself and the other four stages run on 16 registers.
__device__ __forceinline__ void bit_transpose16(uint32_t x[32]) {
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] |= x[k + 16] << 16;
  bit_transpose_stage<8, 0x00FF00FFu, 16>(x);
  bit_transpose_stage<4, 0x0F0F0F0Fu, 16>(x);
  bit_transpose_stage<2, 0x33333333u, 16>(x);
  bit_transpose_stage<1, 0x55555555u, 16>(x);
}

// A lane's plane slices: lift: pl[j], j < 16 the a share and j >= 16 the b
// share of plane j mod 16; MSB: pl[j] a, pl[32 + j] b.
template <bool MSB>
struct PlaneSlices {
  uint32_t pl[MSB ? 64 : 32];
  __device__ __forceinline__ uint32_t a(int j) const { return pl[j]; }
  __device__ __forceinline__ uint32_t b(int j) const { return pl[(MSB ? 32 : 16) + j]; }
};

// Planes when every party's b share is the previous party's a share (the
// reshare's zero-copy views): a wave forms its own a planes only (pl[j], of
// which the lift uses 16) and reads plane j of its b share from the previous
// party's wave through LDS.
struct PlanesPrevB {
  uint32_t pl[32];
  const uint32_t* prev;  // LDS [planes][64] of the previous party, at this lane
  __device__ __forceinline__ uint32_t a(int j) const { return pl[j]; }
  __device__ __forceinline__ uint32_t b(int j) const { return prev[64 * j]; }
};

struct AdderState {
  uint32_t ta, tb;  // plane s - 1
  uint32_t ca, cb;  // carry-save AND of plane s - 1: own, previous party's
  uint32_t xa, xb;  // carry
  uint32_t va, vb;  // last ripple AND output: own, previous party's
};

struct AdderLds {
  // [party][group][writing lane = gate of the batch][its 16 words, padded to 17]: a
  // write of one word by all lanes and a read of one gate by all lanes both
  // spread over the banks (the 17-word row stride staggers them)
  uint32_t rnd[3][4][16][17];
  uint32_t xch[2][3][2][64];   // [step parity][party][ripple, carry-save][lane]
};

// The ChaCha12 block of this lane's gate in batch b (8 steps, 16 gates): gate
// gl of its 16-lane group is step 8 b + gl / 2, the carry-save AND of that
// step's plane (gl even) or its ripple round s - 1 (gl odd; the lift's step 0
// slot is the carry-out). Gates past the circuit (the MSB adder's step 31 and
// its step 0 ripple slot) are not valid: their blocks are never read.
template <bool MSB>
__device__ __forceinline__ void batch_gate(int b, int gl, const RippleCtrs& rc, size_t G, size_t g,
                                           bool& valid, uint64_t& ctr) {
  constexpr int LAST = MSB ? 30 : 15;
  constexpr int LAST_C = MSB ? 29 : 14;
  const int sj = gl >> 1, s = 8 * b + sj;
  // ripple round k = s - 1 of this lane's step: a per-lane index into the
  // counter table (one load from the kernel arguments, not a chain of selects)
  const int k = s - 1;
  const int kc = k < 0 ? (MSB ? 0 : 15) : (k > LAST_C ? LAST_C : k);
  const uint64_t c_round = rc.c[kc];
  if ((gl & 1) == 0) {  // carry-save AND of plane s
    valid = s <= LAST;
    ctr = rc.csa + (uint64_t)s * G + g;
  } else {  // ripple round s - 1; the lift's step 0 slot: carry-out
    valid = (s >= 1 && s <= LAST_C + 1) || (!MSB && s == 0);
    ctr = c_round + g;
  }
}

// The next batch's block, computed a piece per exchange step of the current
// batch so that its VALU work fills the steps' LDS round trips and barriers:
// steps 0-5 run one double round each (column half before the step's barrier,
// diagonal half after it), step 6 adds the input words and hands the block to
// the group through LDS. Step 0's barrier has retired every wave's reads of
// the current batch's blocks before any wave writes the next one, and step
// 7's barrier publishes it.
struct NextDraw {
  uint32_t x[16];
  uint64_t ctr;
  bool valid;
};

template <int J, bool kBeforeBarrier>
__device__ __forceinline__ void draw_piece(NextDraw& d, const ChaChaKey& key, uint64_t nonce,
                                           uint32_t* rnd_row) {
  if constexpr (J < 6) {
    if constexpr (kBeforeBarrier)
      chacha_column_round(d.x);
    else
      chacha_diagonal_round(d.x);
  } else if constexpr (J == 6 && kBeforeBarrier) {
    chacha_finish(key, nonce, d.ctr, d.x);
#pragma unroll
    for (int i = 0; i < 16; i++) rnd_row[i] = d.x[i];  // an invalid gate's block is never read
  }
}

// Step S (last plane LAST). r16[2 (S % 8)] is the carry-save randomness of
// plane S, r16[2 (S % 8) + 1] the ripple randomness of round S - 1; the lift's
// carry-out randomness is r_out. kDraw: this batch computes the next one's
// block (d) between its exchanges.
template <bool MSB, int S, bool kDraw, typename SL>
__device__ __forceinline__ void adder_step(const SL& W, AdderState& st,
                                           const uint32_t r16[16], uint32_t r_out, AdderLds& L,
                                           int p, int prev, int lane, NextDraw& d,
                                           const ChaChaKey& key, uint64_t nonce, uint32_t* rnd_row) {
  constexpr int LAST = MSB ? 30 : 15;
  constexpr int J = S % 8;
  constexpr int X = S & 1;
  if constexpr (S == 0) {
    const uint32_t na = W.a(0), nb = W.b(0);
    const uint32_t cs = csa_and(p, na, nb) ^ r16[0];
    L.xch[X][p][1][lane] = cs;
    if constexpr (kDraw) {
      // the writes go out first, the rounds fill their latency before the barrier
      __builtin_amdgcn_sched_barrier(0);
      draw_piece<J, true>(d, key, nonce, rnd_row);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    st.cb = L.xch[X][prev][1][lane];
    if constexpr (kDraw) draw_piece<J, false>(d, key, nonce, rnd_row);
    st.ca = cs;
    st.ta = na;
    st.tb = nb;
    st.xa = st.xb = st.va = st.vb = 0;
  } else if constexpr (S <= LAST) {
    const uint32_t ba = st.ca ^ (st.ta & party_mask(p, 2)), bb = st.cb ^ (st.tb & party_mask(p, 0));
    const uint32_t na = W.a(S), nb = W.b(S);
    st.xa ^= st.va;  // carry after round S - 2
    st.xb ^= st.vb;
    const uint32_t va = and_pre(na ^ st.xa, nb ^ st.xb, ba ^ st.xa, bb ^ st.xb, r16[2 * J + 1]);
    const uint32_t cs = csa_and(p, na, nb) ^ r16[2 * J];
    L.xch[X][p][0][lane] = va;
    L.xch[X][p][1][lane] = cs;
    if constexpr (kDraw) {
      // the writes go out first, the rounds fill their latency before the barrier
      __builtin_amdgcn_sched_barrier(0);
      draw_piece<J, true>(d, key, nonce, rnd_row);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    st.vb = L.xch[X][prev][0][lane];
    st.cb = L.xch[X][prev][1][lane];
    if constexpr (kDraw) draw_piece<J, false>(d, key, nonce, rnd_row);
    st.va = va;
    st.ca = cs;
    st.ta = na;
    st.tb = nb;
    if constexpr (!MSB && S == LAST) {
      // carry-out (protocol.rs:2229-2239): c2 = b15 & c (shared AND), c1 = c ^ b15
      st.xa ^= st.va;
      st.xb ^= st.vb;
      const uint32_t ba2 = st.ca ^ (st.ta & party_mask(p, 2)), bb2 = st.cb ^ (st.tb & party_mask(p, 0));
      const uint32_t v2 = and_pre(ba2, bb2, st.xa, st.xb, r_out);
      st.xa ^= ba2;
      st.xb ^= bb2;
      L.xch[X ^ 1][p][0][lane] = v2;
      __syncthreads();
      st.vb = L.xch[X ^ 1][prev][0][lane];
      st.va = v2;
    }
  } else {
    static_assert(!kDraw, "only the circuit's last batch has steps past its last plane");
  }
}

template <bool MSB, int B, typename SL>
__device__ __forceinline__ void adder_batch(const SL& W, AdderState& st,
                                            const uint32_t r16[16], uint32_t r_out, AdderLds& L,
                                            int p, int prev, int lane, NextDraw& d,
                                            const ChaChaKey& key, uint64_t nonce, uint32_t* rnd_row) {
  constexpr bool kDraw = B + 1 < (MSB ? 4 : 2);
#define IRISMPC_STEP(j) \
  adder_step<MSB, 8 * B + j, kDraw>(W, st, r16, r_out, L, p, prev, lane, d, key, nonce, rnd_row)
  IRISMPC_STEP(0); IRISMPC_STEP(1); IRISMPC_STEP(2); IRISMPC_STEP(3);
  IRISMPC_STEP(4); IRISMPC_STEP(5); IRISMPC_STEP(6); IRISMPC_STEP(7);
#undef IRISMPC_STEP
}

// The adder's batches of 8 steps, the first batch's blocks already in L.rnd
// (published by a barrier) and d free: each batch reads its gates' bits, then
// runs its steps while computing the next batch's block.
template <bool MSB, typename SL>
__device__ __forceinline__ void adder_batches(const SL& sl, AdderState& st, AdderLds& L, const RippleCtrs& rc,
                                              size_t G, size_t g, int p, int prev, int lane, NextDraw& d,
                                              const ChaChaKey& key, uint64_t nonce) {
  constexpr int N_BATCH = MSB ? 4 : 2;
  const int grp = lane >> 4, gl = lane & 15;
  uint32_t* rnd_row = L.rnd[p][grp][gl];
  uint32_t r_out = 0;
#pragma unroll 1
  for (int b = 0; b < N_BATCH; b++) {
    uint32_t r16[16];  // F(own) ^ F(prev): bits of the 16 gates for this lane's half word
#pragma unroll
    for (int c = 0; c < 16; c++) r16[c] = L.rnd[p][grp][c][gl] ^ L.rnd[prev][grp][c][gl];
    if (!MSB && b == 0) r_out = r16[1];
    if (b + 1 < N_BATCH) {
      batch_gate<MSB>(b + 1, gl, rc, G, g, d.valid, d.ctr);
      chacha_start(key, nonce, d.ctr, d.x);
    }
    switch (b) {
      case 0: adder_batch<MSB, 0>(sl, st, r16, r_out, L, p, prev, lane, d, key, nonce, rnd_row); break;
      case 1: adder_batch<MSB, 1>(sl, st, r16, r_out, L, p, prev, lane, d, key, nonce, rnd_row); break;
      case 2:
        if constexpr (MSB) adder_batch<MSB, 2>(sl, st, r16, r_out, L, p, prev, lane, d, key, nonce, rnd_row);
        break;
      default:
        if constexpr (MSB) adder_batch<MSB, 3>(sl, st, r16, r_out, L, p, prev, lane, d, key, nonce, rnd_row);
        break;
    }
  }
}

// Writes the first batch's block of this lane's gate to L.rnd (published by
// the caller's next barrier).
template <bool MSB>
__device__ __forceinline__ void adder_first_draw(AdderLds& L, const RippleCtrs& rc, size_t G, size_t g, int p,
                                                 int lane, NextDraw& d, const ChaChaKey& key, uint64_t nonce) {
  const int grp = lane >> 4, gl = lane & 15;
  batch_gate<MSB>(0, gl, rc, G, g, d.valid, d.ctr);
  uint32_t k[16] = {};
  if (d.valid) chacha_block<12>(key, nonce, d.ctr, k);
#pragma unroll
  for (int i = 0; i < 16; i++) L.rnd[p][grp][gl][i] = k[i];
}

// The three-launch form's adders (lift: MSB = false; lift_mul_sub and MSB
// adder: MSB = true), for b views that are separate buffers: each wave forms
// the planes of both its shares. (When every b view is the previous party's a
// buffer, compare_fused_kernel runs all stages at once.)
template <bool MSB>
__global__ __launch_bounds__(192) void adder_word_kernel(const Trio t, uint32_t A, const RippleCtrs rc) {
  __shared__ AdderLds L;
  // one party per wave: p is wave-uniform, so its keys and pointers stay scalar
  const int p = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63), prev = (p + 2) % 3;
  const int grp = lane >> 4;  // 16-lane group = 8 words = one ChaCha block
  const CircuitBufs& P = t.p[p];
  const size_t W = t.W, G = W / 8, n_el = 64 * W;
  const size_t hw = (size_t)blockIdx.x * 64 + lane;  // half word: elements 32 hw .. 32 hw + 31
  const size_t g = (size_t)blockIdx.x * 4 + grp;     // ChaCha block of words 8 g .. 8 g + 7
  const bool live = hw < 2 * W;

  // batch 0's block first: its rounds overlap the loads below
  NextDraw d;
  adder_first_draw<MSB>(L, rc, G, g, p, lane, d, P.key_own, t.nonce);

  PlaneSlices<MSB> sl;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const size_t e = 32 * hw + 8 * q;
    if constexpr (MSB) {
      // m = (mask - c1<<16 - c2<<17) * A - code<<16, minus 1 on x_0
      // (kernel.cu:181-195,407-426)
      uint16_t ma[8] = {}, mb[8] = {}, c1a[8] = {}, c2a[8] = {}, c1b[8] = {}, c2b[8] = {},
               xa[8] = {}, xb[8] = {};
      if (live) {
        u16x8_load(P.mask_a + e, ma);
        u16x8_load(P.mask_b + e, mb);
        u16x8_load(P.corr_a + e, c1a);
        u16x8_load(P.corr_a + n_el + e, c2a);
        u16x8_load(t.p[prev].corr_a + e, c1b);  // b share = the previous party's a share
        u16x8_load(t.p[prev].corr_a + n_el + e, c2b);
        u16x8_load(P.code_a + e, xa);
        u16x8_load(P.code_b + e, xb);
      }
      uint32_t la[8], lb[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        la[i] = (uint32_t)ma[i] - ((uint32_t)c1a[i] << 16) - ((uint32_t)c2a[i] << 17);
        lb[i] = (uint32_t)mb[i] - ((uint32_t)c1b[i] << 16) - ((uint32_t)c2b[i] << 17);
      }
      if (P.lifted_a && live) {
        uint4* da = reinterpret_cast<uint4*>(P.lifted_a + e);
        uint4* db = reinterpret_cast<uint4*>(P.lifted_b + e);
        da[0] = uint4{la[0], la[1], la[2], la[3]};
        da[1] = uint4{la[4], la[5], la[6], la[7]};
        db[0] = uint4{lb[0], lb[1], lb[2], lb[3]};
        db[1] = uint4{lb[4], lb[5], lb[6], lb[7]};
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        sl.pl[8 * q + i] = la[i] * A - ((uint32_t)xa[i] << 16) - (p == 0 ? 1u : 0u);
        sl.pl[32 + 8 * q + i] = lb[i] * A - ((uint32_t)xb[i] << 16) - (p == 1 ? 1u : 0u);
      }
    } else {
      uint16_t ma[8] = {}, mb[8] = {};
      if (live) {
        u16x8_load(P.mask_a + e, ma);
        u16x8_load(P.mask_b + e, mb);
      }
#pragma unroll
      for (int i = 0; i < 8; i++) sl.pl[8 * q + i] = (uint32_t)ma[i] | ((uint32_t)mb[i] << 16);
    }
  }
  bit_transpose32(sl.pl);
  if constexpr (MSB) bit_transpose32(sl.pl + 32);
  __syncthreads();

  AdderState st{};
  adder_batches<MSB>(sl, st, L, rc, G, g, p, prev, lane, d, P.key_own, t.nonce);
  if (!live) return;
  if constexpr (MSB) {
    // msb = c ^ a_30 ^ b_30, a_30 = plane 31 (protocol.rs:2347-2352)
    st.xa ^= st.va;
    st.xb ^= st.vb;
    const uint32_t ba = st.ca ^ (st.ta & party_mask(p, 2)), bb = st.cb ^ (st.tb & party_mask(p, 0));
    reinterpret_cast<uint32_t*>(P.res_a)[hw] = st.xa ^ sl.a(31) ^ ba;
    reinterpret_cast<uint32_t*>(P.res_b)[hw] = st.xb ^ sl.b(31) ^ bb;
  } else {
    reinterpret_cast<uint32_t*>(P.c_a)[hw] = st.xa;
    reinterpret_cast<uint32_t*>(P.c_b)[hw] = st.xb;
    reinterpret_cast<uint32_t*>(P.c_a + W)[hw] = st.va;
    reinterpret_cast<uint32_t*>(P.c_b + W)[hw] = st.vb;
  }
}

// ------------------------------------------------------------ fused compare
// lift -> bit injection -> lift_mul_sub -> MSB adder in one launch, for b
// views that are the previous party's a buffers (the scan's zero-copy
// reshare). A block holds the same 2,048 elements (32 words) of the three
// parties, one wave each, and a lane keeps the 32 elements of its half word
// through all the stages:
//  * the lift's carries stay in the lane's registers as bit slices (bit 16:
//    xa, xb; bit 17: va, vb -- the lift kernel's c planes 0 and 1);
//  * the injection of element e draws u16 e of each stream, so a lane's 32
//    elements are exactly ChaCha12 block ctr + hw (bit 16) and ctr + 2 W + hw
//    (bit 17): the lane computes its own party's two blocks and reads the
//    previous party's from LDS, and party 1's messages reach party 2 through
//    LDS;
//  * the corrections enter the lifted mask in registers.
// No carry plane or correction goes through HBM. Shares, counters and opened
// bits are those of the three launches (adder_word_kernel, inject_kernel,
// adder_word_kernel).
__device__ __forceinline__ uint32_t lo16(uint32_t w, int h) { return (w >> (16 * h)) & 0xffffu; }

// bits 2 w and 2 w + 1 of `bits` as the u16 pair (b0, b1)
__device__ __forceinline__ uint32_t pair_bits(uint32_t bits, int w) {
  return ((bits >> (2 * w)) & 1u) | (((bits >> (2 * w + 1)) & 1u) << 16);
}

__global__ __launch_bounds__(192) void compare_fused_kernel(const Trio t, uint32_t A, const RippleCtrs lift_rc,
                                                            uint64_t inject_ctr, const RippleCtrs msb_rc) {
  __shared__ AdderLds L;
  // a planes of the three parties ([party][plane][lane]: lift 16 planes, MSB
  // 32), in between the injection's keystreams ([party][bit][word][lane])
  __shared__ uint32_t xs[3 * 32 * 64];
  const int p = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63), prev = (p + 2) % 3;
  const int grp = lane >> 4;
  const CircuitBufs& P = t.p[p];
  const size_t W = t.W, G = W / 8;
  const size_t hw = (size_t)blockIdx.x * 64 + lane;  // half word: elements 32 hw .. 32 hw + 31
  const size_t g = (size_t)blockIdx.x * 4 + grp;     // adders' ChaCha block of words 8 g .. 8 g + 7
  const bool live = hw < 2 * W;

  // ---- lift (protocol.rs:2041-2074,2137-2241)
  NextDraw d;
  adder_first_draw<false>(L, lift_rc, G, g, p, lane, d, P.key_own, t.nonce);
  uint32_t mk[16];  // the lane's mask a shares, elements 2 i (low half) and 2 i + 1
  PlanesPrevB sl;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint16_t ma[8] = {};
    if (live) u16x8_load(P.mask_a + 32 * hw + 8 * q, ma);
#pragma unroll
    for (int i = 0; i < 8; i++) sl.pl[8 * q + i] = ma[i];
#pragma unroll
    for (int i = 0; i < 4; i++) mk[4 * q + i] = (uint32_t)ma[2 * i] | ((uint32_t)ma[2 * i + 1] << 16);
  }
  bit_transpose16(sl.pl);
#pragma unroll
  for (int j = 0; j < 16; j++) xs[(p * 16 + j) * 64 + lane] = sl.pl[j];
  sl.prev = &xs[prev * 16 * 64 + lane];
  __syncthreads();
  AdderState st{};
  adder_batches<false>(sl, st, L, lift_rc, G, g, p, prev, lane, d, P.key_own, t.nonce);
  const uint32_t bits_a[2] = {st.xa, st.va}, bits_b[2] = {st.xb, st.vb};
  __syncthreads();  // every wave is done with the lift planes (xs) and L.rnd

  // ---- bit injection (protocol.rs:1302-1643, kernel.cu:526-629); messages
  // P1 -> P2 through L.rnd, free between the adders
#pragma unroll
  for (int c = 0; c < 2; c++) {
    uint32_t k[16];
    chacha_block<12>(P.key_own, t.nonce, inject_ctr + (uint64_t)c * 2 * W + hw, k);
#pragma unroll
    for (int w = 0; w < 16; w++) xs[((p * 2 + c) * 16 + w) * 64 + lane] = k[w];
  }
  __syncthreads();
  uint32_t* msg = &L.rnd[0][0][0][0];  // [bit][word][lane], 2,048 of its 3,264 words
  uint32_t cr[2][16];  // corrections (a share) of bits 16 and 17, two elements per word
  // two elements per word throughout: packed u16 arithmetic; bit b of an
  // element as a u16 (bv) and as an all-ones or zero u16 (bm = b * 0xffff)
  if (p != 2) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      // P0 holds its b share of the carry, P1 the XOR of its two shares
      const uint32_t bits = p == 0 ? bits_b[c] : bits_a[c] ^ bits_b[c];
#pragma unroll
      for (int w = 0; w < 16; w++) {
        const u16x2 own = as_u16x2(xs[((p * 2 + c) * 16 + w) * 64 + lane]);
        const u16x2 prv = as_u16x2(xs[((prev * 2 + c) * 16 + w) * 64 + lane]);
        const uint32_t bv = pair_bits(bits, w), bm = bv * 0xffffu;
        if (p == 0) {  // rand_01 = my rng, rand_02 = their rng: y = b r01 - r02, o = r01 - 2 y
          const u16x2 y = (as_u16x2(bm) & own) - prv;
          cr[c][w] = as_u32(own - (y << 1));
        } else {  // rand_01 = their rng, rand_12 = my rng: x = b - r01 (to P2), o = x - 2 r12
          const u16x2 x = as_u16x2(bv) - prv;
          cr[c][w] = as_u32(x - (own << 1));
          msg[(c * 16 + w) * 64 + lane] = as_u32(x);
        }
      }
    }
  }
  __syncthreads();
  if (p == 2) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const uint32_t bits = bits_a[c];
#pragma unroll
      for (int w = 0; w < 16; w++) {
        // rand_12 = their rng, rand_02 = my rng, x = P1's message: z = b x - r12, o = b - 2 (r02 + z)
        const u16x2 own = as_u16x2(xs[((p * 2 + c) * 16 + w) * 64 + lane]);
        const u16x2 prv = as_u16x2(xs[((prev * 2 + c) * 16 + w) * 64 + lane]);
        const u16x2 x = as_u16x2(msg[(c * 16 + w) * 64 + lane]);
        const uint32_t bv = pair_bits(bits, w), bm = bv * 0xffffu;
        const u16x2 z = (as_u16x2(bm) & x) - prv;
        cr[c][w] = as_u32(as_u16x2(bv) - ((own + z) << 1));
      }
    }
  }
  __syncthreads();  // keystreams and messages are read before the MSB stage reuses xs and L.rnd

  // ---- lift_mul_sub (protocol.rs:1831-1872, kernel.cu:181-195,407-426) and MSB adder
  adder_first_draw<true>(L, msb_rc, G, g, p, lane, d, P.key_own, t.nonce);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const size_t e = 32 * hw + 8 * q;
    uint16_t xa[8] = {};
    if (live) u16x8_load(P.code_a + e, xa);
    uint32_t la[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int el = 8 * q + i;
      const uint32_t m = lo16(mk[el >> 1], el & 1);
      const uint32_t c1 = lo16(cr[0][el >> 1], el & 1), c2 = lo16(cr[1][el >> 1], el & 1);
      la[i] = m - (c1 << 16) - (c2 << 17);
      sl.pl[el] = la[i] * A - ((uint32_t)xa[i] << 16) - (p == 0 ? 1u : 0u);
    }
    if (P.lifted_a && live) {  // own a, and the next party's b
      uint4* da = reinterpret_cast<uint4*>(P.lifted_a + e);
      uint4* db = reinterpret_cast<uint4*>(t.p[(p + 1) % 3].lifted_b + e);
      da[0] = db[0] = uint4{la[0], la[1], la[2], la[3]};
      da[1] = db[1] = uint4{la[4], la[5], la[6], la[7]};
    }
  }
  bit_transpose32(sl.pl);
#pragma unroll
  for (int j = 0; j < 32; j++) xs[(p * 32 + j) * 64 + lane] = sl.pl[j];
  sl.prev = &xs[prev * 32 * 64 + lane];
  __syncthreads();
  st = AdderState{};
  adder_batches<true>(sl, st, L, msb_rc, G, g, p, prev, lane, d, P.key_own, t.nonce);
  if (!live) return;
  // msb = c ^ a_30 ^ b_30, a_30 = plane 31 (protocol.rs:2347-2352)
  st.xa ^= st.va;
  st.xb ^= st.vb;
  const uint32_t ba = st.ca ^ (st.ta & party_mask(p, 2)), bb = st.cb ^ (st.tb & party_mask(p, 0));
  reinterpret_cast<uint32_t*>(P.res_a)[hw] = st.xa ^ sl.a(31) ^ ba;
  reinterpret_cast<uint32_t*>(P.res_b)[hw] = st.xb ^ sl.b(31) ^ bb;
}

// ------------------------------------------------------------ open -> bitmap
// Wave per result word: word i covers 64 consecutive DB rows of one query row
// (chunk widths are multiples of 64), so the opened bits land in a 64-bit run
// of bitmap bits starting at totalDb*(q/31) + db + offset: two atomicOr.
__global__ __launch_bounds__(256) void open_kernel(const OpenArgs a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_words) return;
  uint64_t v = a.a[i] ^ a.b[i] ^ a.c[i];
  const size_t e0 = i * 64;
  const size_t q = e0 / a.chunk_len, d0 = e0 % a.chunk_len;
  if (q >= a.query_len || d0 >= a.real_chunk_len) return;
  const size_t valid = a.real_chunk_len - d0;
  if (valid < 64) v &= (1ull << valid) - 1;
  if (!v) return;
  if (a.anon.enabled()) {
    uint64_t m = v;
    while (m) {
      const int k = __builtin_ctzll(m);
      m &= m - 1;
      anon_store(a.anon, a.dots, e0 + (size_t)k, a.offset + d0 + (size_t)k, q);
    }
  }
  const size_t slot = a.slot_of ? a.slot_of[q / kRotations] : q / kRotations;
  const size_t bit = a.total_db_len * slot + d0 + a.offset;
  const size_t wi = bit >> 6, sh = bit & 63;
  atomicOr(reinterpret_cast<unsigned long long*>(&a.bitmap[wi]), (unsigned long long)(v << sh));
  if (sh && (v >> (64 - sh)))
    atomicOr(reinterpret_cast<unsigned long long*>(&a.bitmap[wi + 1]),
             (unsigned long long)(v >> (64 - sh)));
}

// ------------------------------------------------------------ launchers
void launch_adder_word(const Trio& t, bool msb, uint32_t threshold_a, const RippleCtrs& ctrs,
                       hipStream_t s) {
  IRIS_REQUIRE(t.W % 8 == 0, "bit planes must be a multiple of 8 words");
  const unsigned blocks = (unsigned)ceil_div(t.W, 32);
  if (msb)
    adder_word_kernel<true><<<blocks, 192, 0, s>>>(t, threshold_a, ctrs);
  else
    adder_word_kernel<false><<<blocks, 192, 0, s>>>(t, threshold_a, ctrs);
  IRIS_HIP_CHECK(hipGetLastError());
}

bool trio_b_is_prev(const Trio& t) {
  for (int p = 0; p < 3; p++) {
    const int q = (p + 2) % 3;
    if (t.p[p].code_b != t.p[q].code_a || t.p[p].mask_b != t.p[q].mask_a ||
        (t.p[p].lifted_a == nullptr) != (t.p[p].lifted_b == nullptr))
      return false;
  }
  return true;
}

void launch_compare_fused(const Trio& t, uint32_t threshold_a, const RippleCtrs& lift, uint64_t inject_ctr,
                          const RippleCtrs& msb, hipStream_t s) {
  IRIS_REQUIRE(t.W % 8 == 0, "bit planes must be a multiple of 8 words");
  IRIS_REQUIRE(trio_b_is_prev(t), "the fused compare needs b views that are the previous party's a buffers");
  compare_fused_kernel<<<(unsigned)ceil_div(t.W, 32), 192, 0, s>>>(t, threshold_a, lift, inject_ctr, msb);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_inject(const Trio& t, uint64_t ctr, hipStream_t s) {
  inject_kernel<<<(unsigned)ceil_div(128 * t.W, 2048), 192, 0, s>>>(t, ctr);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_open(const OpenArgs& a, hipStream_t s) {
  if (a.n_words == 0) return;
  open_kernel<<<(unsigned)ceil_div(a.n_words, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
