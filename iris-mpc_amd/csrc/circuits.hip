// Bit-sliced replicated-share circuits for gfx950.
//
// Restates threshold_ring/protocol.rs (lift_mpc :2041-2074,
// binary_add_3_get_two_carries :2137-2241, bit_inject :1302-1643,
// lift_mul_sub :1831-1872, extract_msb/binary_add_3_get_msb :2243-2359) and
// threshold_ring/cuda/kernel.cu. Design for one GPU:
//  * the 16x64 / 32x64 bit transposes are wave ballots (lane = element);
//  * x1/x2/x3 of `split` are never materialised: for party p they are fixed
//    selections of the transposed (a, b) planes, so x2^x3, x1^x3 and the
//    carry-save AND collapse to one expression per party (bit-identical);
//  * correlated randomness is generated in registers from the block counter
//    the reference would have used (no keystream buffers);
//  * each launch covers all three parties (grid.y); a party's threads read
//    only that party's state and write messages into the next party's "b"
//    arrays / inbox. The kernel boundary is the exchange barrier.
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

__device__ __forceinline__ int party_of(const Trio& t) { return t.party0 + (int)blockIdx.y; }

// Correlated randomness F(own) ^ F(prev) as 8 u64 words of block `blk`.
__device__ __forceinline__ void corr_u64x8(const CircuitBufs& P, uint64_t nonce, uint64_t blk,
                                           uint64_t r[8]) {
  uint32_t k1[16], k2[16];
  chacha_block<12>(P.key_own, nonce, blk, k1);
  chacha_block<12>(P.key_prev, nonce, blk, k2);
#pragma unroll
  for (int i = 0; i < 8; i++)
    r[i] = (uint64_t)(k1[2 * i] ^ k2[2 * i]) | ((uint64_t)(k1[2 * i + 1] ^ k2[2 * i + 1]) << 32);
}

// 8 consecutive words; every call site addresses a multiple of 8 words from a
// 256-byte aligned allocation (W is a multiple of 8), so 64-byte alignment
// holds and the accesses vectorise.
__device__ __forceinline__ void load8(const uint64_t* p, uint64_t v[8]) {
  const uint64_t* q = static_cast<const uint64_t*>(__builtin_assume_aligned(p, 64));
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = q[i];
}

__device__ __forceinline__ void store8(uint64_t* p, const uint64_t v[8]) {
  uint64_t* q = static_cast<uint64_t*>(__builtin_assume_aligned(p, 64));
#pragma unroll
  for (int i = 0; i < 8; i++) q[i] = v[i];
}

// ------------------------------------------------------------ transpose
// Plane j, word i, bit k = bit j of element 64 i + k (kernel.cu:63-84).
// 64 x 32 bit-matrix transpose across a wave: lane l holds x_l; lanes j and
// 32 + j receive the 64-bit word whose bit l is bit j of x_l. Five butterfly
// stages inside each 32-lane half swap the off-diagonal s x s blocks with the
// partner lane l ^ s (ds_swizzle): a per-lane rotate of the partner's value
// and one bitfield insert. The halves are then joined across lane ^ 32.
template <int S, uint32_t LOW>
__device__ __forceinline__ uint32_t transpose_stage(uint32_t x, int lane) {
  const uint32_t y = (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (S << 10));
  const bool hi = (lane & S) != 0;
  const uint32_t keep = hi ? ~LOW : LOW;
  const uint32_t rot = __builtin_rotateright32(y, hi ? S : 32 - S);
  return (x & keep) | (rot & ~keep);
}

__device__ __forceinline__ uint64_t wave_transpose32(uint32_t x, int lane) {
  x = transpose_stage<16, 0x0000FFFFu>(x, lane);
  x = transpose_stage<8, 0x00FF00FFu>(x, lane);
  x = transpose_stage<4, 0x0F0F0F0Fu>(x, lane);
  x = transpose_stage<2, 0x33333333u>(x, lane);
  x = transpose_stage<1, 0x55555555u>(x, lane);
  const uint32_t other = (uint32_t)__shfl_xor((int)x, 32);
  return lane < 32 ? ((uint64_t)x | ((uint64_t)other << 32)) : ((uint64_t)other | ((uint64_t)x << 32));
}

// Block per 64 consecutive words: each wave turns 16 words into bit planes
// (a in bits 0..15 and b in bits 16..31 of one transpose), the planes go
// through LDS and leave as 512-byte runs per plane.
__global__ __launch_bounds__(256) void transpose16_kernel(const Trio t) {
  const int p = party_of(t);
  const CircuitBufs& P = t.p[p];
  __shared__ uint64_t planes[32][64];  // planes 0..15: a, 16..31: b
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t w0 = (size_t)blockIdx.x * 64;
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const int wl = wave * 16 + k;
    const size_t w = w0 + wl;
    uint32_t va = 0, vb = 0;
    if (w < t.W) {
      va = P.mask_a[w * 64 + lane];
      vb = P.mask_b[w * 64 + lane];
    }
    const uint64_t plane = wave_transpose32(va | (vb << 16), lane);
    if (lane < 32) planes[lane][wl] = plane;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const int idx = (int)threadIdx.x + 256 * m, pl = idx >> 6, wl = idx & 63;
    const size_t w = w0 + wl;
    if (w >= t.W) continue;
    if (pl < 16)
      P.tm_a[(size_t)pl * t.W + w] = planes[pl][wl];
    else
      P.tm_b[(size_t)(pl - 16) * t.W + w] = planes[pl][wl];
  }
}

// ------------------------------------------------------------ carry-save AND
// and_pre(x1^x3, x2^x3) for the split of party p (kernel.cu:39-42,205-235):
//   p0: x1=(a,0) x2=0 x3=(0,b)  -> a & b
//   p1: x1=(0,b) x2=(a,0) x3=0  -> a & b
//   p2: x1=0 x2=(0,b) x3=(a,0)  -> a ^ (a & b)
__device__ __forceinline__ uint64_t csa_and(int p, uint64_t a, uint64_t b) {
  return p == 2 ? (a & ~b) : (a & b);
}

// ------------------------------------------------------------ ripple helpers
// a_k = s plane k = transposed plane k+1; b_k = carry-save plane k ^ x3 plane k
// (x3 = (0, b) for p0, (a, 0) for p2, 0 for p1).
__device__ __forceinline__ uint64_t and_pre(uint64_t xa, uint64_t xb, uint64_t ya, uint64_t yb,
                                            uint64_t r) {
  return (xa & ya) ^ (xb & ya) ^ (xa & yb) ^ r;
}

// ------------------------------------------------------------ fused adder
// The carry-save stage and every ripple round of one adder (the lift: 16
// planes, first + 14 rounds + carry-out; the MSB: 31 planes, first + 29
// rounds + final XOR) in one launch. Wave p of a block is party p; lane l
// handles words 8 (64 blockIdx.x + l) .. + 7. Each round loads one
// transposed plane pair, T_{k+1}: it is a_k of this round and the input of
// the carry-save AND of plane k + 1, whose output b_{k+1} the next round uses.
// Both AND outputs of a round reach the next party through one LDS exchange;
// carry-save outputs and the running carry never leave the block.
__device__ __forceinline__ void csa8(int p, const uint64_t a[8], const uint64_t b[8],
                                     const uint64_t r[8], uint64_t v[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = csa_and(p, a[i], b[i]) ^ r[i];
}

__global__ __launch_bounds__(192) void adder_fused_kernel(const Trio t, int msb, const RippleCtrs rc) {
  __shared__ uint64_t xbuf[2][3][64][16];  // per round slot: ripple AND (8), carry-save AND (8)
  const int p = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63), prev = (p + 2) % 3;
  const CircuitBufs& P = t.p[p];
  const size_t W = t.W, g = (size_t)blockIdx.x * 64 + lane, o = 8 * g, G = W / 8;
  const bool live = o < W;
  const uint64_t* Ta = msb ? P.tx_a : P.tm_a;
  const uint64_t* Tb = msb ? P.tx_b : P.tm_b;
  const int last = msb ? 30 : 15;  // ripple rounds 0 .. last - 1, carry planes 0 .. last
  uint64_t ta[8] = {}, tb[8] = {}, ca[8] = {}, cb[8] = {}, xa[8] = {}, xb[8] = {}, va[8] = {},
           vb[8] = {}, r[8];

  // carry-save AND of plane 0
  if (live) {
    load8(Ta + o, ta);
    load8(Tb + o, tb);
    corr_u64x8(P, t.nonce, rc.csa + g, r);
    csa8(p, ta, tb, r, ca);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) xbuf[1][p][lane][8 + i] = ca[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; i++) cb[i] = xbuf[1][prev][lane][8 + i];

  for (int k = 0; k < last; k++) {
    uint64_t na[8] = {}, nb[8] = {}, cs[8] = {};
    // b_k = carry-save plane k ^ x3 (x3 = (0, Tb_k) for p0, (Ta_k, 0) for p2)
    uint64_t ba[8], bb[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      ba[i] = ca[i] ^ (p == 2 ? ta[i] : 0);
      bb[i] = cb[i] ^ (p == 0 ? tb[i] : 0);
    }
    if (live) {
      load8(Ta + (size_t)(k + 1) * W + o, na);  // a_k = T plane k + 1
      load8(Tb + (size_t)(k + 1) * W + o, nb);
      corr_u64x8(P, t.nonce, rc.c[k] + g, r);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      xa[i] ^= va[i];  // carry after round k - 1 (va = vb = 0 before round 1)
      xb[i] ^= vb[i];
      va[i] = live ? and_pre(na[i] ^ xa[i], nb[i] ^ xb[i], ba[i] ^ xa[i], bb[i] ^ xb[i], r[i]) : 0;
    }
    if (k == 0) {  // the first round has no carry in
#pragma unroll
      for (int i = 0; i < 8; i++) {
        xa[i] = 0;
        xb[i] = 0;
      }
    }
    if (live) {  // carry-save AND of plane k + 1 (planes 1 .. last)
      corr_u64x8(P, t.nonce, rc.csa + (size_t)(k + 1) * G + g, r);
      csa8(p, na, nb, r, cs);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      xbuf[k & 1][p][lane][i] = va[i];
      xbuf[k & 1][p][lane][8 + i] = cs[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
      vb[i] = xbuf[k & 1][prev][lane][i];
      cb[i] = xbuf[k & 1][prev][lane][8 + i];
      ca[i] = cs[i];
      ta[i] = na[i];
      tb[i] = nb[i];
    }
  }
  // carry after the last round
#pragma unroll
  for (int i = 0; i < 8; i++) {
    xa[i] ^= va[i];
    xb[i] ^= vb[i];
  }
  // b_last = carry-save plane last ^ x3
  uint64_t ba[8], bb[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    ba[i] = ca[i] ^ (p == 2 ? ta[i] : 0);
    bb[i] = cb[i] ^ (p == 0 ? tb[i] : 0);
  }
  if (!msb) {
    // carry-out (protocol.rs:2229-2239): c2 = b15 & c (shared AND), c1 = c ^ b15
    if (live) corr_u64x8(P, t.nonce, rc.c[15] + g, r);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      va[i] = live ? and_pre(ba[i], bb[i], xa[i], xb[i], r[i]) : 0;
      xa[i] ^= ba[i];
      xb[i] ^= bb[i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) xbuf[(last - 1 + 1) & 1][p][lane][i] = va[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) vb[i] = xbuf[(last - 1 + 1) & 1][prev][lane][i];
    if (live) {
      store8(P.c_a + o, xa);
      store8(P.c_b + o, xb);
      store8(P.c_a + W + o, va);
      store8(P.c_b + W + o, vb);
    }
    return;
  }
  // MSB (protocol.rs:2347-2352): msb = c ^ a_30 ^ b_30, a_30 = T plane 31
  if (!live) return;
  uint64_t aa[8], ab[8];
  load8(Ta + 31 * W + o, aa);
  load8(Tb + 31 * W + o, ab);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    xa[i] ^= aa[i] ^ ba[i];
    xb[i] ^= ab[i] ^ bb[i];
  }
  store8(P.res_a + o, xa);
  store8(P.res_b + o, xb);
}

// ------------------------------------------------------------ bit injection
// 128 W elements e: bits of the two lift carries (e < 64 W: bit 16,
// e >= 64 W: bit 17); ChaCha12 block g of each stream covers elements 32 g .. 32 g + 31.
// kernel.cu:526-629 with the message flow P1 -> P2 -> P0 and P0 -> P1.

// Wave per 2048 elements (64 ChaCha12 blocks, one per lane; block g covers
// elements 32 g .. 32 g + 31 as the u16 view of its keystream). The two
// keystreams go through LDS in element order, then each lane handles 8
// consecutive elements per pass with 16-byte loads and stores.
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

__global__ __launch_bounds__(256) void inject_kernel(const Trio t, int step, uint64_t ctr) {
  const int p = party_of(t);
  // which party acts in which step
  const bool act = (step == 1 && p != 2) || (step == 2 && p != 0) || (step == 3 && p == 0);
  if (!act) return;
  const CircuitBufs& P = t.p[p];
  const CircuitBufs& N = t.p[(p + 1) % 3];
  __shared__ uint32_t ks[4][2][1024];  // [wave][own, prev][u16 pairs in element order]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t n = 128 * t.W;
  const size_t group = (size_t)blockIdx.x * 4 + wave;
  const size_t e_base = group * 2048;
  if (e_base < n) {
    uint32_t ko[16], kp[16];
    chacha_block<12>(P.key_own, t.nonce, ctr + group * 64 + lane, ko);   // my_rng (chacha1)
    chacha_block<12>(P.key_prev, t.nonce, ctr + group * 64 + lane, kp);  // their_rng (chacha2)
#pragma unroll
    for (int i = 0; i < 16; i += 4) {
      *reinterpret_cast<uint4*>(&ks[wave][0][lane * 16 + i]) = uint4{ko[i], ko[i + 1], ko[i + 2], ko[i + 3]};
      *reinterpret_cast<uint4*>(&ks[wave][1][lane * 16 + i]) = uint4{kp[i], kp[i + 1], kp[i + 2], kp[i + 3]};
    }
  }
  __syncthreads();
  if (e_base >= n) return;
  const uint16_t* own16 = reinterpret_cast<const uint16_t*>(ks[wave][0]);
  const uint16_t* prv16 = reinterpret_cast<const uint16_t*>(ks[wave][1]);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int el = 8 * (k * 64 + lane);
    const size_t e = e_base + el;
    if (e >= n) break;
    uint16_t own[8], prv[8], in[8], o1[8], o2[8], o3[8];
    u16x8_load(own16 + el, own);
    u16x8_load(prv16 + el, prv);
    if (p == 0) {
      // rand_01 = my rng, rand_02 = their rng; in_b = party 0's b share
      const uint32_t bits = (uint32_t)(P.c_b[e >> 6] >> (e & 63)) & 0xffu;
      if (step == 1) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint16_t b = (uint16_t)((bits >> i) & 1u), r01 = own[i], r02 = prv[i];
          const uint16_t y = (uint16_t)(b * r01 - r02);
          o1[i] = y;
          o2[i] = (uint16_t)(r01 - (y << 1));
        }
        u16x8_store(N.inbox + e, o1);
        u16x8_store(P.corr_a + e, o2);
      } else {
        u16x8_load(P.inbox + e, in);
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint16_t b = (uint16_t)((bits >> i) & 1u), r02 = prv[i], z = in[i];
          o1[i] = (uint16_t)(-((r02 + z) << 1) + b);
        }
        u16x8_store(P.corr_b + e, o1);
      }
    } else if (p == 1) {
      // rand_01 = their rng, rand_12 = my rng
      if (step == 1) {
        const uint32_t bits = (uint32_t)((P.c_a[e >> 6] ^ P.c_b[e >> 6]) >> (e & 63)) & 0xffu;
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const uint16_t r01 = prv[i], r12 = own[i];
          const uint16_t x = (uint16_t)(((bits >> i) & 1u) - r01);
          o1[i] = x;
          o2[i] = (uint16_t)(x - (r12 << 1));
        }
        u16x8_store(N.inbox + e, o1);
        u16x8_store(P.corr_a + e, o2);
      } else {
        u16x8_load(P.inbox + e, in);
#pragma unroll
        for (int i = 0; i < 8; i++) o1[i] = (uint16_t)(prv[i] - (in[i] << 1));
        u16x8_store(P.corr_b + e, o1);
      }
    } else {
      // rand_12 = their rng, rand_02 = my rng; in_a = party 2's a share
      const uint32_t bits = (uint32_t)(P.c_a[e >> 6] >> (e & 63)) & 0xffu;
      u16x8_load(P.inbox + e, in);
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const uint16_t r12 = prv[i], r02 = own[i], x = in[i], b = (uint16_t)((bits >> i) & 1u);
        const uint16_t z = (uint16_t)(b * x - r12);
        o1[i] = z;
        o2[i] = (uint16_t)(-((r02 + z) << 1) + b);
        o3[i] = (uint16_t)(x - (r12 << 1));
      }
      u16x8_store(N.inbox + e, o1);
      u16x8_store(P.corr_a + e, o2);
      u16x8_store(P.corr_b + e, o3);
    }
  }
}

// ------------------------------------------------------------ lift_mul_sub + transpose32
// m = (mask - c1<<16 - c2<<17) * A - code<<16, minus 1 on x_0 (kernel.cu:181-195,
// 407-426), then the 32x64 bit transpose (kernel.cu:86-107).
__global__ __launch_bounds__(256) void lms_transpose32_kernel(const Trio t, uint32_t A) {
  const int p = party_of(t);
  const CircuitBufs& P = t.p[p];
  __shared__ uint64_t planes[64][64];  // planes 0..31: a, 32..63: b
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const size_t w0 = (size_t)blockIdx.x * 64;
  const size_t L = 64 * t.W;
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const int wl = wave * 16 + k;
    const size_t w = w0 + wl;
    uint32_t ma = 0, mb = 0;
    if (w < t.W) {
      const size_t e = w * 64 + lane;
      ma = (uint32_t)P.mask_a[e] - ((uint32_t)P.corr_a[e] << 16) - ((uint32_t)P.corr_a[e + L] << 17);
      mb = (uint32_t)P.mask_b[e] - ((uint32_t)P.corr_b[e] << 16) - ((uint32_t)P.corr_b[e + L] << 17);
      if (P.lifted_a) {
        P.lifted_a[e] = ma;
        P.lifted_b[e] = mb;
      }
      ma = ma * A - ((uint32_t)P.code_a[e] << 16);
      mb = mb * A - ((uint32_t)P.code_b[e] << 16);
      if (p == 0) ma -= 1u;
      if (p == 1) mb -= 1u;
    }
    // plane j of a in lane j, of b in lane 32 + j
    const uint64_t pa = wave_transpose32(ma, lane), pb = wave_transpose32(mb, lane);
    planes[lane][wl] = lane < 32 ? pa : pb;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int idx = (int)threadIdx.x + 256 * m, pl = idx >> 6, wl = idx & 63;
    const size_t w = w0 + wl;
    if (w >= t.W) continue;
    if (pl < 32)
      P.tx_a[(size_t)pl * t.W + w] = planes[pl][wl];
    else
      P.tx_b[(size_t)(pl - 32) * t.W + w] = planes[pl][wl];
  }
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
void launch_transpose16(const Trio& t, int np, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(t.W, 64), (unsigned)np);
  transpose16_kernel<<<grid, 256, 0, s>>>(t);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_adder_fused(const Trio& t, bool msb, const RippleCtrs& ctrs, hipStream_t s) {
  IRIS_REQUIRE(t.party0 == 0, "the fused adder runs all three parties");
  IRIS_REQUIRE(t.W % 8 == 0, "bit planes must be a multiple of 8 words");
  adder_fused_kernel<<<(unsigned)ceil_div(t.W, 512), 192, 0, s>>>(t, msb ? 1 : 0, ctrs);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_inject(const Trio& t, int np, int step, uint64_t ctr, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(128 * t.W, 8192), (unsigned)np);
  inject_kernel<<<grid, 256, 0, s>>>(t, step, ctr);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_lms_transpose32(const Trio& t, int np, uint32_t A, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(t.W, 64), (unsigned)np);
  lms_transpose32_kernel<<<grid, 256, 0, s>>>(t, A);
  IRIS_HIP_CHECK(hipGetLastError());
}

void launch_open(const OpenArgs& a, hipStream_t s) {
  if (a.n_words == 0) return;
  open_kernel<<<(unsigned)ceil_div(a.n_words, 256), 256, 0, s>>>(a);
  IRIS_HIP_CHECK(hipGetLastError());
}

}  // namespace irismpc
