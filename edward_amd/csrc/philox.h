// Counter-based RNG for momenta and accept uniforms.
//
// Philox4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): a bijection of a 128-bit counter keyed
// by 64 bits.  Each (chain, transition) pair owns its own counters, so draws do
// not depend on launch shape, chain count or call history, and a resumed run
// continues bit-identically.
#pragma once

#include "common.h"

namespace edhmc {

struct u32x4 { uint32_t x, y, z, w; };

__device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Stream layout: ctr = (a, chain, counter_lo, counter_hi | stream << 31).
__device__ inline u32x4 philox_draw(uint64_t seed, uint64_t counter,
                                    uint32_t a, uint32_t chain, int stream) {
  u32x4 c;
  c.x = a;
  c.y = chain;
  c.z = (uint32_t)(counter & 0xffffffffu);
  c.w = (uint32_t)((counter >> 32) & 0x7fffffffu) | (stream ? 0x80000000u : 0u);
  return philox4x32_10(c, (uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32));
}

// Box-Muller on two 32-bit words: u1 = ((x0 >> 8) + 1) 2^-24 in (0, 1],
// theta = 2 pi (x1 >> 8) 2^-24.
__device__ inline void box_muller(uint32_t x0, uint32_t x1, float& n0, float& n1) {
  float u1 = (float)((x0 >> 8) + 1u) * 5.9604644775390625e-8f;
  float th = (float)(x1 >> 8) * (5.9604644775390625e-8f * 6.283185307179586f);
  float rad = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincosf(th, &s, &c);
  n0 = rad * c;
  n1 = rad * s;
}

// Four N(0,1) draws for coordinates 4j..4j+3 of `chain` at transition `counter`.
__device__ inline f32x4_t momentum4(uint64_t seed, uint64_t counter, uint32_t j,
                                    uint32_t chain) {
  u32x4 x = philox_draw(seed, counter, j, chain, 0);
  f32x4_t r;
  float a, b;
  box_muller(x.x, x.y, a, b);
  r[0] = a; r[1] = b;
  box_muller(x.z, x.w, a, b);
  r[2] = a; r[3] = b;
  return r;
}

// Accept uniform in [0, 1) with 53 random bits.
__device__ inline double accept_uniform(uint64_t seed, uint64_t counter, uint32_t chain) {
  u32x4 x = philox_draw(seed, counter, 0u, chain, 1);
  uint64_t m = ((uint64_t)x.x << 21) | (uint64_t)(x.y >> 11);
  return (double)m * 1.1102230246251565e-16;  // 2^-53
}

}  // namespace edhmc
