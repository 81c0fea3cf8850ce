/* Generated by gen_mul_sched.py: lane schedules of the short product in mul.
 * Per lane one 32-bit word: bits 0-4 alpha of segment 1's first round, 5-9 its beta,
 * 10-14 / 15-19 the same for segment 2, 20-24 / 25-29 the corner's guard pair
 * (alpha, beta), bit 30 set when the lane has a corner. */
#ifndef FPA_MUL_SCHED_H
#define FPA_MUL_SCHED_H

#include <stdint.h>

namespace fpa {

template <int L> struct MulSched { static constexpr bool ok = false; };

/* L = 4: 2 + 1 rounds per lane (12 slots for 10 full pairs), 2 guard pairs whole, 1 corners; 7 lane-chunk collisions over 5 emissions */
__constant__ uint32_t k_mul_sched_4[4] = {
  0x44010460u, 0x00000821u, 0x00000c42u, 0x00008c62u,
};
template <> struct MulSched<4> {
  static constexpr bool ok = true;
  static constexpr int R1 = 2, R2 = 1, NCORNER = 1;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_4[gl]; }
};

/* L = 8: 3 + 2 rounds per lane (40 slots for 36 full pairs), 4 guard pairs whole, 3 corners; 14 lane-chunk collisions over 7 emissions */
__constant__ uint32_t k_mul_sched_8[8] = {
  0x482304e0u, 0x4a120ca2u, 0x4c011063u, 0x000018c3u, 0x000118e3u, 0x00019825u, 0x00031885u, 0x000398a5u,
};
template <> struct MulSched<8> {
  static constexpr bool ok = true;
  static constexpr int R1 = 3, R2 = 2, NCORNER = 3;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_8[gl]; }
};

/* L = 16: 7 + 3 rounds per lane (160 slots for 136 full pairs), 13 guard pairs whole, 2 corners; 38 lane-chunk collisions over 12 emissions */
__constant__ uint32_t k_mul_sched_16[16] = {
  0x44c71dc0u, 0x5a179de0u, 0x0001a922u, 0x00022942u, 0x00072962u, 0x0007a982u, 0x000035a2u, 0x0000b4c6u,
  0x000134e6u, 0x0001b506u, 0x000234a9u, 0x00033529u, 0x0003b549u, 0x00043569u, 0x00073589u, 0x0007b5a9u,
};
template <> struct MulSched<16> {
  static constexpr bool ok = true;
  static constexpr int R1 = 7, R2 = 3, NCORNER = 2;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_16[gl]; }
};

/* L = 32: 13 + 6 rounds per lane (608 slots for 528 full pairs), 27 guard pairs whole, 4 corners; 74 lane-chunk collisions over 21 emissions */
__constant__ uint32_t k_mul_sched_32[32] = {
  0x4b96bb40u, 0x59273b60u, 0x64c35380u, 0x7253d3a0u, 0x000453c0u, 0x0004d3e0u, 0x00055266u, 0x0006d286u,
  0x000752a6u, 0x00006ac6u, 0x0000eae6u, 0x00016b06u, 0x0001eb26u, 0x000269edu, 0x0002ea0du, 0x00036a2du,
  0x0003ea4du, 0x00046b4du, 0x0004eb6du, 0x00056b8du, 0x0006ebadu, 0x00076bcdu, 0x0007ebedu, 0x00086973u,
  0x0008e993u, 0x00096a73u, 0x000d6a93u, 0x000deab3u, 0x000e6ad3u, 0x000eeaf3u, 0x000f6b13u, 0x000feb33u,
};
template <> struct MulSched<32> {
  static constexpr bool ok = true;
  static constexpr int R1 = 13, R2 = 6, NCORNER = 4;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32[gl]; }
};

}  // namespace fpa

#endif
