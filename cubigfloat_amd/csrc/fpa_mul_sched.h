/* Generated by gen_mul_sched.py: lane schedules of the short product in mul.
 * Per lane one 32-bit word: bits 0-4 alpha of segment 1's first round, 5-9 its beta,
 * 10-14 / 15-19 the same for segment 2, 20-24 / 25-29 the corner's guard pair
 * (alpha, beta), bit 30 set when the lane has a corner.  MulSched<L, F0>: groups of L
 * lanes, full chunk pairs on anti-diagonals >= F0, guard pairs on F0 - 1. */
#ifndef FPA_MUL_SCHED_H
#define FPA_MUL_SCHED_H

#include <stdint.h>

namespace fpa {

template <int L, int F0> struct MulSched { static constexpr bool ok = false; };

/* L = 4, F0 = 3: 2 + 1 rounds per lane (12 slots for 10 full pairs), 2 guard pairs whole, 1 corners; 7 lane-chunk collisions over 5 emissions */
__constant__ uint32_t k_mul_sched_4_3[4] = {
  0x44010460u, 0x00000821u, 0x00000c42u, 0x00008c62u,
};
template <> struct MulSched<4, 3> {
  static constexpr bool ok = true;
  static constexpr int R1 = 2, R2 = 1, NCORNER = 1;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_4_3[gl]; }
};

/* L = 8, F0 = 7: 3 + 2 rounds per lane (40 slots for 36 full pairs), 4 guard pairs whole, 3 corners; 14 lane-chunk collisions over 7 emissions */
__constant__ uint32_t k_mul_sched_8_7[8] = {
  0x482304e0u, 0x4a120ca2u, 0x4c011063u, 0x000018c3u, 0x000118e3u, 0x00019825u, 0x00031885u, 0x000398a5u,
};
template <> struct MulSched<8, 7> {
  static constexpr bool ok = true;
  static constexpr int R1 = 3, R2 = 2, NCORNER = 3;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_8_7[gl]; }
};

/* L = 16, F0 = 15: 7 + 3 rounds per lane (160 slots for 136 full pairs), 13 guard pairs whole, 2 corners; 38 lane-chunk collisions over 12 emissions */
__constant__ uint32_t k_mul_sched_16_15[16] = {
  0x44c71dc0u, 0x5a179de0u, 0x0001a922u, 0x00022942u, 0x00072962u, 0x0007a982u, 0x000035a2u, 0x0000b4c6u,
  0x000134e6u, 0x0001b506u, 0x000234a9u, 0x00033529u, 0x0003b549u, 0x00043569u, 0x00073589u, 0x0007b5a9u,
};
template <> struct MulSched<16, 15> {
  static constexpr bool ok = true;
  static constexpr int R1 = 7, R2 = 3, NCORNER = 2;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_16_15[gl]; }
};

/* L = 32, F0 = 31: 13 + 6 rounds per lane (608 slots for 528 full pairs), 27 guard pairs whole, 4 corners; 74 lane-chunk collisions over 21 emissions */
__constant__ uint32_t k_mul_sched_32_31[32] = {
  0x4b96bb40u, 0x59273b60u, 0x64c35380u, 0x7253d3a0u, 0x000453c0u, 0x0004d3e0u, 0x00055266u, 0x0006d286u,
  0x000752a6u, 0x00006ac6u, 0x0000eae6u, 0x00016b06u, 0x0001eb26u, 0x000269edu, 0x0002ea0du, 0x00036a2du,
  0x0003ea4du, 0x00046b4du, 0x0004eb6du, 0x00056b8du, 0x0006ebadu, 0x00076bcdu, 0x0007ebedu, 0x00086973u,
  0x0008e993u, 0x00096a73u, 0x000d6a93u, 0x000deab3u, 0x000e6ad3u, 0x000eeaf3u, 0x000f6b13u, 0x000feb33u,
};
template <> struct MulSched<32, 31> {
  static constexpr bool ok = true;
  static constexpr int R1 = 13, R2 = 6, NCORNER = 4;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32_31[gl]; }
};

/* L = 32, F0 = 30: 14 + 6 rounds per lane (640 slots for 559 full pairs), 26 guard pairs whole, 4 corners; 74 lane-chunk collisions over 22 emissions */
__constant__ uint32_t k_mul_sched_32_30[32] = {
  0x499dbb60u, 0x591e3b80u, 0x64bebba0u, 0x743f3bc0u, 0x000fbbe0u, 0x0002d264u, 0x00035284u, 0x0003d2a4u,
  0x000dd2c4u, 0x000e52e4u, 0x000ed304u, 0x000f5324u, 0x000fd344u, 0x000069acu, 0x0000e9ccu, 0x000169ecu,
  0x0001ea0cu, 0x00026a2cu, 0x0002ea4cu, 0x00036912u, 0x0003e932u, 0x0006e952u, 0x00076972u, 0x0007e992u,
  0x00086a72u, 0x0008ea92u, 0x00096ab2u, 0x000dead2u, 0x000e6af2u, 0x000eeb12u, 0x000f6b32u, 0x000feb52u,
};
template <> struct MulSched<32, 30> {
  static constexpr bool ok = true;
  static constexpr int R1 = 14, R2 = 6, NCORNER = 4;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32_30[gl]; }
};

/* L = 32, F0 = 29: 12 + 8 rounds per lane (640 slots for 589 full pairs), 23 guard pairs whole, 6 corners; 83 lane-chunk collisions over 22 emissions */
__constant__ uint32_t k_mul_sched_32_29[32] = {
  0x4b754340u, 0x5335c360u, 0x5af64380u, 0x62b6c3a0u, 0x6a7b43c0u, 0x723bc3e0u, 0x000c42c4u, 0x000cc2e4u,
  0x00006304u, 0x0000e324u, 0x00016248u, 0x0001e268u, 0x00026288u, 0x0002e2a8u, 0x000561ccu, 0x0005e1ecu,
  0x0006620cu, 0x0006e22cu, 0x0007634cu, 0x0007e36cu, 0x0008638cu, 0x0008e3acu, 0x000b63ccu, 0x000be3ecu,
  0x000c60d4u, 0x000ce0f4u, 0x000d6114u, 0x000de134u, 0x000e6254u, 0x000ee274u, 0x000f6294u, 0x000fe2b4u,
};
template <> struct MulSched<32, 29> {
  static constexpr bool ok = true;
  static constexpr int R1 = 12, R2 = 8, NCORNER = 6;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32_29[gl]; }
};

/* L = 32, F0 = 28: 16 + 6 rounds per lane (704 slots for 618 full pairs), 28 guard pairs whole, 0 corners; 76 lane-chunk collisions over 24 emissions */
__constant__ uint32_t k_mul_sched_32_28[32] = {
  0x00092280u, 0x0009a2a0u, 0x00063ac0u, 0x0006bae0u, 0x00093b00u, 0x0009bb20u, 0x00015340u, 0x0001d360u,
  0x00025380u, 0x0002d3a0u, 0x000353c0u, 0x0003d3e0u, 0x000651cau, 0x0006d1eau, 0x0009520au, 0x0009d22au,
  0x00006910u, 0x0000e930u, 0x00016950u, 0x0001e970u, 0x00026a90u, 0x0002eab0u, 0x00036ad0u, 0x0003eaf0u,
  0x00066b10u, 0x0006eb30u, 0x00076b50u, 0x0007eb70u, 0x00086b90u, 0x0008ebb0u, 0x00096bd0u, 0x0009ebf0u,
};
template <> struct MulSched<32, 28> {
  static constexpr bool ok = true;
  static constexpr int R1 = 16, R2 = 6, NCORNER = 0;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32_28[gl]; }
};

/* L = 32, F0 = 27: 16 + 6 rounds per lane (704 slots for 646 full pairs), 24 guard pairs whole, 3 corners; 94 lane-chunk collisions over 24 emissions */
__constant__ uint32_t k_mul_sched_32_27[32] = {
  0x439922e0u, 0x56f63b00u, 0x62993b20u, 0x00015340u, 0x0001d360u, 0x00025380u, 0x0002d3a0u, 0x000353c0u,
  0x000653e0u, 0x00095264u, 0x0009d284u, 0x000a52a4u, 0x000ad2c4u, 0x000b51aau, 0x000069cau, 0x0000e9eau,
  0x00016a0au, 0x0001ea2au, 0x000268f0u, 0x0002e910u, 0x00036930u, 0x00066950u, 0x0006e970u, 0x00076af0u,
  0x0007eb10u, 0x00086b30u, 0x0008eb50u, 0x00096b70u, 0x0009eb90u, 0x000a6bb0u, 0x000aebd0u, 0x000b6bf0u,
};
template <> struct MulSched<32, 27> {
  static constexpr bool ok = true;
  static constexpr int R1 = 16, R2 = 6, NCORNER = 3;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32_27[gl]; }
};

/* L = 32, F0 = 23: 16 + 10 rounds per lane (832 slots for 748 full pairs), 23 guard pairs whole, 0 corners; 55 lane-chunk collisions over 28 emissions */
__constant__ uint32_t k_mul_sched_32_23[32] = {
  0x00068a80u, 0x00070aa0u, 0x00078ac0u, 0x00080ae0u, 0x00088b00u, 0x00090b20u, 0x00098b40u, 0x0003b360u,
  0x00043380u, 0x0004b3a0u, 0x000533c0u, 0x0006b3e0u, 0x00073166u, 0x0007b186u, 0x00083030u, 0x0008b050u,
  0x00093070u, 0x0009b090u, 0x000058b0u, 0x0003d8d0u, 0x00045a90u, 0x0004dab0u, 0x00055ad0u, 0x0005daf0u,
  0x00065b10u, 0x0006db30u, 0x00075b50u, 0x0007db70u, 0x00085b90u, 0x0008dbb0u, 0x00095bd0u, 0x0009dbf0u,
};
template <> struct MulSched<32, 23> {
  static constexpr bool ok = true;
  static constexpr int R1 = 16, R2 = 10, NCORNER = 0;
  __device__ static uint32_t lane(int gl) { return k_mul_sched_32_23[gl]; }
};

}  // namespace fpa

#endif
