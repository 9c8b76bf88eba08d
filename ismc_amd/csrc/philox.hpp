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
