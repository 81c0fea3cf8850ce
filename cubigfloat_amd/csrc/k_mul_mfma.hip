/*
 * 4096-bit fpa_mul on the matrix cores (v_mfma_i32_32x32x32_i8), one value per wave.
 * Same results as the radix-2^27 kernel (fpa_warp_mul.h:14-139 semantics): the exact
 * 8192-bit product, the same normalisation, round word and sticky, then round_mantissa.
 *
 * Product as a byte convolution.  The 512-byte mantissas are recoded to signed digits:
 * t = x + 0x8080..80 (one group add with ballot carries), digit_i = byte_i(t) - 128, so
 * x = x_s + c*X with X = 2^4096 and c the add's carry out (1 for any value with its top
 * bit set).  Every digit fits int8 and every column sum c_k = sum_{i+j=k} a_i b_j of the
 * signed digits has |c_k| < 512 * 2^14, so the i32 accumulator is exact.  With k = 32n+m:
 *     C[m][n] = sum_{d=0..16} T_d[m][kk] B_d[kk][n],
 *     T_d[m][kk] = b_s[32d + m - kk]   (a Toeplitz block of b's digits, zero outside b),
 *     B_d[kk][n] = a_s[32(n-d) + kk]   (a's digit chunk n-d, zero outside [0, 16)),
 * so 17 MFMAs accumulate all 1024 columns in one 32x32 tile.  T_d rows are 16 bytes of b's
 * digits in descending order: b is kept reversed in LDS in four byte-shifted copies, so
 * each lane reads its row as four aligned dwords.  B_{d+1} is B_d moved up one lane
 * (DPP wave_shr:1): lane 0 receives 0, and lane 32 receives lane 31's chunk 31-d, which
 * is past a's top for every d <= 15.
 * Then a*b = sum_k c_k 256^k + X*(ca*b_s + cb*a_s)  (mod X^2, and 0 <= a*b < X^2): the
 * high-half term is signed digits too, added by two more MFMAs with an identity T (a's
 * chunks at columns 16..31 are B_16 itself; b's come from LDS).  Column groups of four go
 * through LDS so that lane L owns product limbs 4L..4L+3; signed carries between lanes
 * are resolved by the ballot lookahead.  Lanes 32..63 then hold the top half and round
 * it as a 32-lane group with 4 limbs per lane.
 */
#include "fpa_kernel_common.h"

namespace fpa {

typedef int mm_v4 __attribute__((ext_vector_type(4)));
typedef int mm_v16 __attribute__((ext_vector_type(16)));

/* per-wave LDS, bytes: four reversed copies of b's digits, then a's and b's digits.  A
 * copy holds 144 dwords (dwords 8..135 = b reversed, zeros around) in a 168-dword slot,
 * which puts the four copies 8 banks apart.  The column stage (256 x 8 bytes, element w
 * at w + w/16) reuses [0, 2176). */
constexpr int MM_TS = 672;
constexpr int MM_A = 4 * MM_TS;
constexpr int MM_B = MM_A + 512;
constexpr int MM_WAVE = MM_B + 512;

__device__ __forceinline__ uint32_t mm_wave_shr1(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x138, 0xF, 0xF, true);
}

__global__ __launch_bounds__(256) void fpa_mul_mfma_kernel(fpa_arguments a) {
  __shared__ __attribute__((aligned(16))) uint8_t smem[4 * MM_WAVE];
  const int lane = (int)__lane_id(), wib = threadIdx.x >> 6;
  const uint32_t v = blockIdx.x * 4 + wib;
  if (v >= a.count) return;
  uint8_t *ws = smem + wib * MM_WAVE;
  const uint32_t ea = a.a_exps[v], eb = a.b_exps[v];
  const uint32_t act = k_mul_tab[fpa_classify(ea) * 9 + fpa_classify(eb)];
  const Group<32> half;                /* lanes 32..63: the result, 4 limbs per lane */
  Slot s;
  s.v = v;
  s.valid = lane >= 32;
  if (act != FPA_A_MUL) {              /* uniform over the wave */
    store_special<4>(a, s, half.gl, act);
    return;
  }
  const Group<64> g;
  uint32_t A[2], B[2];
  load_chunk(A, a.a_limbs + (size_t)v * 128 + 2 * lane);
  load_chunk(B, a.b_limbs + (size_t)v * 128 + 2 * lane);

  /* signed digits: t = x + 0x8080..80 */
  const uint32_t KC[2] = {0x80808080u, 0x80808080u};
  uint32_t TA[2], TB[2];
  const uint32_t ca = g.add(TA, A, KC);
  const uint32_t cb = g.add(TB, B, KC);

  /* LDS: a's digits; reversed b digits R (R dword 135-w = bswap(b digit dword w)) in four
   * copies R_s[z] = R[z + s]; dwords 0..7 and 136..143 of R are zero */
  *reinterpret_cast<uint2 *>(ws + MM_A + 8 * lane) = make_uint2(TA[0] ^ 0x80808080u, TA[1] ^ 0x80808080u);
  *reinterpret_cast<uint2 *>(ws + MM_B + 8 * lane) = make_uint2(TB[0] ^ 0x80808080u, TB[1] ^ 0x80808080u);
  {
    const int sc = lane >> 4, k = lane & 15;
    *reinterpret_cast<uint32_t *>(ws + sc * MM_TS + 4 * (k < 8 ? k : 128 + k)) = 0u;
  }
  const uint32_t r_hi = __builtin_bswap32(TB[0] ^ 0x80808080u);   /* R dword 135 - 2L */
  const uint32_t r_lo = __builtin_bswap32(TB[1] ^ 0x80808080u);   /* R dword 134 - 2L */
  const uint32_t r_up = mm_wave_shr1(r_lo);                        /* R dword 136 - 2L (0 in lane 0) */
  *reinterpret_cast<uint2 *>(ws + 4 * (134 - 2 * lane)) = make_uint2(r_lo, r_hi);
#pragma unroll
  for (int sh = 1; sh < 4; sh++) {
    *reinterpret_cast<uint2 *>(ws + sh * MM_TS + 4 * (134 - 2 * lane)) =
        make_uint2(__builtin_amdgcn_alignbyte(r_hi, r_lo, sh), __builtin_amdgcn_alignbyte(r_up, r_hi, sh));
    if (lane == 63) *reinterpret_cast<uint32_t *>(ws + sh * MM_TS + 28) = __builtin_amdgcn_alignbyte(r_lo, 0u, sh);
  }

  /* 17 MFMAs: lane (h, r) holds row r of T_d and column r of B_d, bytes 16h..16h+15 of k */
  const int h = lane >> 5, r = lane & 31;
  mm_v4 bb = {0, 0, 0, 0}, bs = {0, 0, 0, 0};
  if (r < 16) bb = *reinterpret_cast<const mm_v4 *>(ws + MM_A + 32 * r + 16 * h);
  else bs = *reinterpret_cast<const mm_v4 *>(ws + MM_B + 32 * (r - 16) + 16 * h);   /* b's chunk r-16 */
  const uint32_t sl = (3u - (uint32_t)r) & 3u;                     /* byte shift of this row */
  int qb = (int)((543u - (uint32_t)r + 16u * (uint32_t)h - sl) >> 2) - 128 + (int)sl * (MM_TS / 4);   /* >= 0 */
  asm volatile("" : "+v"(qb));         /* one base register; the 17 rows are immediate offsets */
  const uint32_t *tb = reinterpret_cast<const uint32_t *>(ws) + qb;
  mm_v16 acc = {};
#pragma unroll
  for (int d = 0; d <= 16; d++) {
    mm_v4 t;
    t.x = (int)tb[128 - 8 * d];
    t.y = (int)tb[129 - 8 * d];
    t.z = (int)tb[130 - 8 * d];
    t.w = (int)tb[131 - 8 * d];
    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(t, bb, acc, 0, 0, 0);
    if (d < 16) {
      bb.x = (int)mm_wave_shr1((uint32_t)bb.x);
      bb.y = (int)mm_wave_shr1((uint32_t)bb.y);
      bb.z = (int)mm_wave_shr1((uint32_t)bb.z);
      bb.w = (int)mm_wave_shr1((uint32_t)bb.w);
    }
  }

  /* high half: + cb*a_s and + ca*b_s at columns 512.. (T = identity: row m has a 1 at k = m) */
  {
    const int j = r - 16 * h;          /* byte of this lane's 16 that is the diagonal */
    mm_v4 ti = {0, 0, 0, 0};
    if (j >= 0 && j < 16) {
      const int one = 1 << (8 * (j & 3));
      ti.x = (j >> 2) == 0 ? one : 0;
      ti.y = (j >> 2) == 1 ? one : 0;
      ti.z = (j >> 2) == 2 ? one : 0;
      ti.w = (j >> 2) == 3 ? one : 0;
    }
    if (cb) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(ti, bb, acc, 0, 0, 0);   /* bb = B_16 */
    if (ca) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(ti, bs, acc, 0, 0, 0);
  }

  /* acc[4q+i] = column 32r + 8q + 4h + i: dword w = 8r + 2q + h of the product gets
   * V_w = sum_i acc[4q+i] 2^(8i) (signed, |V_w| < 2^48), as three 32x32->64 signed mads */
  int64_t *V = reinterpret_cast<int64_t *>(ws);
  int w8 = 1 << 8, w16 = 1 << 16, w24 = 1 << 24;
  asm volatile("" : "+s"(w8), "+s"(w16), "+s"(w24));   /* multiplies, not 64-bit shifts */
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int64_t vq = (int64_t)acc[4 * q];
    vq = (int64_t)acc[4 * q + 1] * (int64_t)w8 + vq;
    vq = (int64_t)acc[4 * q + 2] * (int64_t)w16 + vq;
    vq = (int64_t)acc[4 * q + 3] * (int64_t)w24 + vq;
    V[8 * r + 2 * q + h + (r >> 1)] = vq;   /* element w at w + w/16 */
  }
  /* lane L: product limbs 4L..4L+3 = low word of V_w plus the high part of V_(w-1) */
  const int base = 4 * lane + (lane >> 2);
  int32_t hprev = 0;
  if (lane) hprev = reinterpret_cast<const int32_t *>(V)[2 * (4 * lane - 1 + ((lane - 1) >> 2)) + 1];
  uint32_t D[4];
  int64_t c = 0;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int64_t vt = V[base + t];
    const int64_t sum = (int64_t)(uint32_t)vt + (int64_t)hprev + c;
    D[t] = (uint32_t)sum;
    c = sum >> 32;                     /* in [-1, 1] */
    hprev = (int32_t)(vt >> 32);
  }
  /* signed carry from the lane below, then the 1-bit carries/borrows it leaves; the carry
   * out of limb 255 is the X^2 term and is dropped */
  const int ci = (int)g.from_prev((uint32_t)(int32_t)c);
  g.add_lane_words(D, ci > 0 ? (uint32_t)ci : 0u);
  g.sub_lane_words(D, ci < 0 ? (uint32_t)(-ci) : 0u);

  /* P in [2^8190, 2^8192) for normal operands: if the top bit is clear, shift left once */
  const bool full = (__builtin_amdgcn_readlane(D[3], 63) >> 31) != 0;
  if (!full) {
    const uint32_t pv = g.from_prev(D[3]);
    D[3] = __builtin_amdgcn_alignbit(D[3], D[2], 31);
    D[2] = __builtin_amdgcn_alignbit(D[2], D[1], 31);
    D[1] = __builtin_amdgcn_alignbit(D[1], D[0], 31);
    D[0] = __builtin_amdgcn_alignbit(D[0], pv, 31);
  }
  /* mantissa = limbs 128..255 (lanes 32..63), round word = limb 127, sticky = limbs 0..126 */
  const uint32_t rw = __builtin_amdgcn_readlane(D[3], 31);
  const uint32_t loc = lane < 31 ? (D[0] | D[1] | D[2] | D[3]) : lane == 31 ? (D[0] | D[1] | D[2]) : 0u;
  const uint32_t sticky = __ballot(loc != 0) != 0 ? 1u : 0u;
  const uint32_t sign = (ea ^ eb) & 1u;
  int64_t E = (int64_t)(ea >> 1) + (int64_t)(eb >> 1) - FPA_BIAS_LL - (full ? 0 : 1);
  if (round_mantissa<4, 32>(half, D, (int)a.mode, sign, rw, sticky)) E += 1;
  store_result(a, s, half.gl, pack_exp(E, sign, 5, 0x7FFFFFFFll), D);
}

}  // namespace fpa

/* 4096-bit mul on the matrix cores; -1 for any other precision */
extern "C" int fpa_impl_launch_mul_mfma(const fpa_arguments *a, void *stream) {
  if (a->precision != 4096u) return -1;
  if (a->count == 0) return 0;
  hipLaunchKernelGGL(fpa::fpa_mul_mfma_kernel, dim3((a->count + 3u) / 4u), dim3(256), 0, (hipStream_t)stream, *a);
  return (int)hipGetLastError();
}
