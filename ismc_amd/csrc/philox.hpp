// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11), the one counter-based generator of libismc:
// the transport kernel draws a packet's uniforms from it and the lookup
// table kernel its normals, so both streams follow this single definition
// (and the oracle's restatement, which the Random123 known-answer vectors
// pin).
//
// Ten rounds; per round the two 32x32->64 products are one v_mad_u64_u32
// each, and each new word's three-input XOR is one v_bitop3_b32 (truth
// table 0x96, three-input parity; the compiler does not form it from two
// v_xor_b32 by itself).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ismc_rng {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;  // key schedule (golden ratio)
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;  // key schedule (sqrt 3 - 1)

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// counter (c0, c1, c2, c3), key (k0, k1) -> out[0..3]
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
    const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

}  // namespace ismc_rng

namespace ismc_rng {

// The same Philox4x32-10 on counters (c0, ev, c2, c3) of which only c1 = ev
// changes from block to block (a packet's event index; c0, c2, c3 name the
// packet and its stream).  With the changing word in c1, which the first
// round does not multiply, four of the first three rounds' six products
// and the key mixing around them depend on the packet alone:
//   round 1: M0 c0 and M1 c2;  round 2: M1 (round 1's c2);  round 3: M0
//   (round 2's c0).
// PhiloxFixed holds what they contribute, taken once per packet; a block
// then costs 16 v_mad_u64_u32 instead of 20.  philox_block(fixed(...), ev)
// equals philox4x32_10(c0, ev, c2, c3) bit for bit.
struct PhiloxFixed {
  uint32_t a;  // round 1: hi(M1 c2) ^ k0, so that c0' = a ^ ev
  uint32_t d;  // round 1: lo(M0 c0), round 2's c3
  uint32_t f;  // round 2: lo(M1 c2'), round 3's c1
  uint32_t g;  // round 3: hi(M0 c0'') ^ k1'', so that c2''' = g ^ c3''
  uint32_t h;  // round 3: lo(M0 c0''), round 4's c3
};

__device__ __forceinline__ PhiloxFixed philox_fixed(uint32_t c0, uint32_t c2, uint32_t c3, uint32_t k0,
                                                    uint32_t k1) {
  PhiloxFixed f;
  // round 1
  const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
  const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
  f.a = (uint32_t)(p1 >> 32) ^ k0;
  const uint32_t b1 = (uint32_t)p1;                                // c1 after round 1
  const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);           // c2 after round 1
  f.d = (uint32_t)p0;                                              // c3 after round 1
  // round 2: the product of c2
  const uint64_t q1 = (uint64_t)kPhiloxM1 * n2;
  const uint32_t e = xor3((uint32_t)(q1 >> 32), b1, k0 + kPhiloxW0);  // c0 after round 2
  f.f = (uint32_t)q1;                                              // c1 after round 2
  // round 3: the product of c0
  const uint64_t r0 = (uint64_t)kPhiloxM0 * e;
  f.g = (uint32_t)(r0 >> 32) ^ (k1 + 2u * kPhiloxW1);
  f.h = (uint32_t)r0;                                              // c3 after round 3
  return f;
}

__device__ __forceinline__ void philox_block(const PhiloxFixed& f, uint32_t ev, uint32_t k0, uint32_t k1,
                                             uint32_t out[4]) {
  // round 1: c0 = a ^ ev (c1, c2, c3 fixed)
  const uint32_t c0r1 = f.a ^ ev;
  // round 2: the product of c0
  const uint64_t p0 = (uint64_t)kPhiloxM0 * c0r1;
  const uint32_t c2r2 = xor3((uint32_t)(p0 >> 32), f.d, k1 + kPhiloxW1);
  const uint32_t c3r2 = (uint32_t)p0;
  // round 3: the product of c2
  const uint64_t p1 = (uint64_t)kPhiloxM1 * c2r2;
  uint32_t c0 = xor3((uint32_t)(p1 >> 32), f.f, k0 + 2u * kPhiloxW0);
  uint32_t c1 = (uint32_t)p1;
  uint32_t c2 = f.g ^ c3r2;
  uint32_t c3 = f.h;
  // rounds 4 .. 10
  k0 += 3u * kPhiloxW0;
  k1 += 3u * kPhiloxW1;
#pragma unroll
  for (int r = 3; r < 10; ++r) {
    const uint64_t q0 = (uint64_t)kPhiloxM0 * c0;
    const uint64_t q1 = (uint64_t)kPhiloxM1 * c2;
    const uint32_t n0 = xor3((uint32_t)(q1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(q0 >> 32), c3, k1);
    c1 = (uint32_t)q1;
    c3 = (uint32_t)q0;
    c0 = n0;
    c2 = n2;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

}  // namespace ismc_rng
