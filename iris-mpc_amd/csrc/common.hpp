// Shared host/device definitions for the irismpc HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace irismpc {

constexpr int kCodeLen = 12800;   // IRIS_CODE_LENGTH (iris-mpc-common/src/lib.rs:15)
constexpr int kMaskLen = 6400;    // MASK_CODE_LENGTH (lib.rs:16)
constexpr int kRotations = 31;    // ROTATIONS (lib.rs:17)
constexpr int kDbChunk = 1 << 15; // DB_CHUNK_SIZE (iris-mpc-gpu/src/server/actor.rs:115)
constexpr int kAllMatchesLen = 256;  // MAX_MATCHES_LEN (dot/kernel.cu:5)

// Tiled limb layout: a "fragment" is 16 rows x 64 coefficients of one limb,
// stored as 64 lanes x 16 bytes so one wave loads it with a single
// coalesced 1 KiB dwordx4 access straight into MFMA operand registers.
// Lane l holds row (l % 16), coefficients 16*(l/16) .. +15 of the k-tile.
constexpr int kRowsPerGroup = 16;
constexpr int kKTile = 64;
constexpr int kFragBytes = 1024;

struct Error : std::runtime_error {
  explicit Error(const std::string& s) : std::runtime_error(s) {}
};

#define IRIS_HIP_CHECK(expr)                                                          \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      throw ::irismpc::Error(std::string(#expr) + " failed: " + hipGetErrorString(_e) + \
                             " (" __FILE__ ":" + std::to_string(__LINE__) + ")");       \
  } while (0)

#define IRIS_REQUIRE(cond, msg)                                                              \
  do {                                                                                       \
    if (!(cond)) throw ::irismpc::Error(std::string("precondition failed: ") + (msg));        \
  } while (0)

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t round_up(size_t a, size_t b) { return ceil_div(a, b) * b; }

// ChaCha key material passed to kernels by value.
struct ChaChaKey {
  uint32_t k[8];
};

// Two u16 elements of a word, for packed 16-bit arithmetic (v_pk_*_u16).
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_u16x2(uint32_t w) { return __builtin_bit_cast(u16x2, w); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

}  // namespace irismpc
