// u1_philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), shared by host
// and gfx950 code.  It replaces the reference's per-site cuRAND MRG32k3a states
// (random.h:31-58, random.cu:16-36): a draw is a pure function of
// (key = seed, counter = (link, sweep, stream|pass|call, replica)), so there is
// no state array to read and write on every Metropolis launch, and a run can be
// resumed from (configuration, seed, sweep index) alone.
//
// Counter layout used on this path (all four words are 32 bit):
//   c0 = link index        eo_id + parity*V/2 + mu*V       (or site index)
//   c1 = sweep / step index
//   c2 = stream << 24 | pass << 12 | call      (call = hit pair, or draw slot)
//   c3 = replica index (update stream) or measurement-call index (other streams)
// Each Philox call yields four 32-bit words; a Metropolis hit consumes two
// (proposal, accept test), so one call serves two hits.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define U1_HD __host__ __device__ inline
#else
#define U1_HD inline
#endif

namespace u1r {

enum Stream : uint32_t {
  STREAM_UPDATE = 0,     // Metropolis of the main chain
  STREAM_HOTSTART = 1,   // random initial angles
  STREAM_MULTIHIT = 2,   // stochastic multihit averages
  STREAM_ML_L4 = 3,      // multilevel, level-4 sublattice updates
  STREAM_ML_L2 = 4,      // multilevel, level-2 sublattice updates
  STREAM_ML_MULTIHIT = 5 // multilevel, stochastic multihit
};

struct U32x4 {
  uint32_t v[4];
};

static constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
static constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
static constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;  // golden ratio
static constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;  // sqrt(3) - 1

// a ^ b ^ c.  On gfx950 one v_bitop3_b32 (truth table 0x96) instead of two
// v_xor_b32; the compiler does not form it from the XOR chain by itself.
U1_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// One Philox4x32 round with round key (k0, k1).  FUSED selects xor3 for the
// two three-way XORs.
template <bool FUSED>
U1_HD U32x4 philox_round(U32x4 c, uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)PHILOX_M0 * c.v[0];
  const uint64_t p1 = (uint64_t)PHILOX_M1 * c.v[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  U32x4 n;
  n.v[0] = FUSED ? xor3(hi1, c.v[1], k0) : hi1 ^ c.v[1] ^ k0;
  n.v[1] = lo1;
  n.v[2] = FUSED ? xor3(hi0, c.v[3], k1) : hi0 ^ c.v[3] ^ k1;
  n.v[3] = lo0;
  return n;
}

U1_HD U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
  // Round 1 keeps plain XORs: with the link index in c0 and wavefront-uniform
  // c1..c3 and key, its first output word is uniform and stays in scalar
  // registers (a vector bitop3 would move it to the vector unit).
  c = philox_round<false>(c, k0, k1);
#pragma unroll
  for (int r = 1; r < 10; ++r) {
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
    c = philox_round<true>(c, k0, k1);
  }
  return c;
}

U1_HD U32x4 draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
  U32x4 c;
  c.v[0] = c0; c.v[1] = c1; c.v[2] = c2; c.v[3] = c3;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

U1_HD uint32_t ctr2(uint32_t stream, uint32_t pass, uint32_t call) {
  return (stream << 24) | ((pass & 0xFFFu) << 12) | (call & 0xFFFu);
}

// Uniform in [0, 1): the FP64 path uses all 32 bits (exact in binary64), the
// FP32 path the top 24 bits (exact in binary32).  Both read the same word, so
// the two precisions see the same random numbers up to 2^-24.
U1_HD double u01_d(uint32_t w) { return (double)w * 0x1p-32; }
U1_HD float u01_f(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }

}  // namespace u1r
