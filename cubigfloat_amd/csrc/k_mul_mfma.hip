/*
 * fpa_mul on the matrix cores (v_mfma_i32_32x32x32_i8), one value per wave, for the
 * precisions whose limb count n is a multiple of 64 (2048, 4096, 6144, 8192, 16384 bits).
 * Same results as the radix-2^27 kernel (fpa_warp_mul.h:14-139 semantics): the exact
 * 2n-limb product, the same normalisation, round word and sticky, then round_mantissa.
 *
 * Product as a byte convolution.  The 4n-byte mantissas are recoded to signed digits:
 * t = x + 0x8080..80 (one group add with ballot carries), digit_i = byte_i(t) - 128, so
 * x = x_s + c*X with X = 2^(32n) and c the add's carry out (1 for any value with its top
 * bit set).  Every digit fits int8 and every column sum c_k = sum_{i+j=k} a_i b_j of the
 * signed digits has |c_k| < 4n * 2^14 <= 2^25, so the i32 accumulators are exact.  With
 * NCH = n/8 chunks of 32 bytes and k = 32N + m (N = 32t + n the column of tile t):
 *     C_t[m][n] = sum_{d=0..NCH} T_d[m][kk] B_(d,t)[kk][n],
 *     T_d[m][kk] = b_s[32d + m - kk]          (a Toeplitz block of b's digits),
 *     B_(d,t)[kk][n] = a_s[32(32t + n - d) + kk]  (a's chunk 32t+n-d, zero outside [0, NCH)),
 * one MFMA per (d, t) whose tile meets a's chunks.  T_d rows are 16 bytes of b's digits in
 * descending order: b is kept reversed in LDS in four byte-shifted copies, so each lane
 * reads its row as four aligned dwords; each T_d serves every tile.  With one tile
 * (NCH <= 16) B_(d+1) is B_d moved up one lane (DPP wave_shr:1): lane 0 receives 0, and
 * lane 32 receives lane 31's chunk 31-d, which is past a's top for every d < NCH.  With
 * several tiles each B_(d,t) is one 16-byte LDS read from a zero-padded copy of a's chunks.
 * Then a*b = sum_k c_k 256^k + X*(ca*b_s + cb*a_s)  (mod X^2, and 0 <= a*b < X^2): the
 * high-half term is signed digits too, added by MFMAs with an identity T (a's chunks at
 * columns NCH.. are B_(NCH,t); b's come from LDS).  Column groups of four go through LDS
 * so that lane L owns product limbs KP*L .. KP*L+KP-1 (KP = n/32); signed carries between
 * lanes are resolved by the ballot lookahead.  Lanes 32..63 then hold the top half and
 * round it as a 32-lane group with KP limbs per lane.
 */
#include "fpa_kernel_common.h"

namespace fpa {

typedef int mm_v4 __attribute__((ext_vector_type(4)));
typedef int mm_v16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t mm_wave_shr1(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x138, 0xF, 0xF, true);
}

/* Per-wave LDS layout (bytes) for NCH chunks of 32 bytes.  R: four reversed copies of b's
 * digits, each 8*NCH+16 dwords (dwords 8..8NCH+7 = b reversed, zeros around) in a slot of
 * RS dwords, RS = 8 mod 32 so that the copies sit 8 banks apart.  BD: b's digits.  With
 * one tile, AD: a's digits; with several, AP: a's chunks in two 16-byte halves, chunk c at
 * 16*(c+32) of half h, zeros for c outside [0, NCH).  The column stage (2n x 8 bytes,
 * element w at w + w/16) reuses the start of R. */
template <int NCH>
struct MmLayout {
  static constexpr int NT = (NCH + 15) / 16;                 /* 32-column tiles of the product */
  static constexpr int RS = 32 * ((8 * NCH + 8 + 31) / 32) + 8;
  static constexpr int R = 0;
  static constexpr int BD = R + 16 * RS;
  static constexpr int AD = BD + 32 * NCH;                    /* one tile */
  static constexpr int AP = AD;                               /* several tiles */
  static constexpr int APH = 16 * (NCH + 64);                 /* one half of AP */
  static constexpr int WAVE = NT > 1 ? AP + 2 * APH : AD + 32 * NCH;
  static constexpr int STAGE = 8 * (16 * NCH + NCH);          /* column stage bytes */
  static_assert(STAGE <= BD, "column stage fits in the R copies");
};

template <int NCH>
__global__ __launch_bounds__(256) void fpa_mul_mfma_kernel(fpa_arguments a) {
  typedef MmLayout<NCH> Lay;
  constexpr int n = 8 * NCH;           /* limbs */
  constexpr int KI = n / 64;           /* operand limbs per lane */
  constexpr int KP = n / 32;           /* product limbs per lane */
  constexpr int NT = Lay::NT;
  static_assert(n % 64 == 0, "one wave holds the operands lane-contiguously");
  __shared__ __attribute__((aligned(16))) uint8_t smem[4 * Lay::WAVE];
  const int lane = (int)__lane_id(), wib = threadIdx.x >> 6;
  const uint32_t v = blockIdx.x * 4 + wib;
  if (v >= a.count) return;
  uint8_t *ws = smem + wib * Lay::WAVE;
  const uint32_t ea = a.a_exps[v], eb = a.b_exps[v];
  const uint32_t act = k_mul_tab[fpa_classify(ea) * 9 + fpa_classify(eb)];
  const Group<32> half;                /* lanes 32..63: the result, KP limbs per lane */
  Slot s;
  s.v = v;
  s.valid = lane >= 32;
  if (act != FPA_A_MUL) {              /* uniform over the wave */
    store_special<KP>(a, s, half.gl, act);
    return;
  }
  const Group<64> g;
  uint32_t A[KI], B[KI];
  load_chunk(A, a.a_limbs + (size_t)v * n + KI * lane);
  load_chunk(B, a.b_limbs + (size_t)v * n + KI * lane);

  /* signed digits: t = x + 0x8080..80 */
  uint32_t KC[KI];
#pragma unroll
  for (int i = 0; i < KI; i++) KC[i] = 0x80808080u;
  uint32_t TA[KI], TB[KI];
  const uint32_t ca = g.add(TA, A, KC);
  const uint32_t cb = g.add(TB, B, KC);
#pragma unroll
  for (int i = 0; i < KI; i++) { TA[i] ^= 0x80808080u; TB[i] ^= 0x80808080u; }

  /* LDS: digits of a and b; reversed b, R dword j = bswap(b digit dword 8NCH+7-j), in four
   * copies R_s[z] = R[z + s]; dwords 0..7 and 8NCH+8..8NCH+15 of R are zero */
  uint32_t *w32 = reinterpret_cast<uint32_t *>(ws);
#pragma unroll
  for (int i = 0; i < KI; i++) {
    if constexpr (NT == 1) w32[Lay::AD / 4 + KI * lane + i] = TA[i];
    w32[Lay::BD / 4 + KI * lane + i] = TB[i];
  }
  {
    const int sc = lane >> 4, k = lane & 15;
    w32[sc * Lay::RS + (k < 8 ? k : 8 * NCH + k)] = 0u;
  }
  uint32_t RD[KI];                     /* RD[i] = R dword 8NCH+7 - (KI*lane + i) */
#pragma unroll
  for (int i = 0; i < KI; i++) RD[i] = __builtin_bswap32(TB[i]);
  const uint32_t r_up = mm_wave_shr1(RD[KI - 1]);   /* R dword 8NCH+8 - KI*lane (0 in lane 0) */
  const int jlo = 8 * NCH + 8 - KI * (lane + 1);     /* this lane's lowest R dword */
#pragma unroll
  for (int sh = 0; sh < 4; sh++) {
#pragma unroll
    for (int i = 0; i < KI; i++) {
      /* R_s dword jlo + (KI-1-i) from R dwords jlo + KI-1-i (= RD[i]) and the one above */
      const uint32_t lo = RD[i], hi = i == 0 ? r_up : RD[i - 1];
      w32[sh * Lay::RS + jlo + (KI - 1 - i)] = sh == 0 ? lo : __builtin_amdgcn_alignbyte(hi, lo, sh);
    }
    if (sh > 0 && lane == 63) w32[sh * Lay::RS + 7] = __builtin_amdgcn_alignbyte(RD[KI - 1], 0u, sh);
  }
  const int h = lane >> 5, r = lane & 31;
  if constexpr (NT > 1) {
    const int c = lane & 31;           /* zero chunks: 32 below a and 32 above it, both halves */
    uint4 z = {0, 0, 0, 0};
    *reinterpret_cast<uint4 *>(ws + Lay::AP + Lay::APH * h + 16 * c) = z;
    *reinterpret_cast<uint4 *>(ws + Lay::AP + Lay::APH * h + 16 * (NCH + 32 + c)) = z;
#pragma unroll
    for (int i = 0; i < KI; i++) {     /* a's digit dword w = KI*lane+i: chunk w/8, half (w/4)&1 */
      const int w = KI * lane + i;
      w32[(Lay::AP + Lay::APH * ((w >> 2) & 1) + 16 * ((w >> 3) + 32)) / 4 + (w & 3)] = TA[i];
    }
  }

  /* MFMAs: lane (h, r) holds row r of T_d and column r of each B, bytes 16h..16h+15 of k */
  const uint32_t sl = (3u - (uint32_t)r) & 3u;                     /* byte shift of this row */
  int qb = (int)((32u * NCH + 31u - (uint32_t)r + 16u * (uint32_t)h - sl) >> 2) - 8 * NCH +
           (int)sl * Lay::RS;                                      /* T_NCH's first dword, >= 0 */
  asm volatile("" : "+v"(qb));         /* one base register; the rows are immediate offsets */
  const uint32_t *tb = w32 + qb;
  mm_v16 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = mm_v16{};
  mm_v4 bb = {0, 0, 0, 0};
  if constexpr (NT == 1) {
    if (r < NCH) bb = *reinterpret_cast<const mm_v4 *>(ws + Lay::AD + 32 * r + 16 * h);
  }
  const uint8_t *aph = ws + Lay::AP + Lay::APH * h;
#pragma unroll
  for (int d = 0; d <= NCH; d++) {
    const int o = 8 * (NCH - d);
    mm_v4 t4;
    t4.x = (int)tb[o];
    t4.y = (int)tb[o + 1];
    t4.z = (int)tb[o + 2];
    t4.w = (int)tb[o + 3];
    if constexpr (NT == 1) {
      acc[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(t4, bb, acc[0], 0, 0, 0);
      if (d < NCH) {
        bb.x = (int)mm_wave_shr1((uint32_t)bb.x);
        bb.y = (int)mm_wave_shr1((uint32_t)bb.y);
        bb.z = (int)mm_wave_shr1((uint32_t)bb.z);
        bb.w = (int)mm_wave_shr1((uint32_t)bb.w);
      }
    } else {
#pragma unroll
      for (int t = 0; t < NT; t++) {
        /* tile t meets a's chunks when [32t - d, 32t + 31 - d] overlaps [0, NCH) */
        if (32 * t + 31 - d >= 0 && 32 * t - d <= NCH - 1) {
          const mm_v4 bt = *reinterpret_cast<const mm_v4 *>(aph + 16 * (32 * t + r - d + 32));
          acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(t4, bt, acc[t], 0, 0, 0);
        }
      }
      /* keep each step's operand reads within reach of its MFMAs: hoisting them across the
       * whole unrolled loop holds hundreds of registers */
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  /* high half: + cb*a_s and + ca*b_s at chunk columns NCH..2NCH-1 (T = identity: row m of
   * T has a 1 at k = m) */
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
#pragma unroll
    for (int t = 0; t < NT; t++) {
      if (32 * t + 31 < NCH) continue;                 /* tile below the high half */
      const int c = 32 * t + r - NCH;                  /* digit chunk at column 32t + r */
      mm_v4 ba, bs = {0, 0, 0, 0};
      if constexpr (NT == 1) ba = bb;                  /* B_NCH */
      else ba = *reinterpret_cast<const mm_v4 *>(aph + 16 * (c + 32));
      if (c >= 0 && c < NCH) bs = *reinterpret_cast<const mm_v4 *>(ws + Lay::BD + 32 * c + 16 * h);
      if (cb) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ti, ba, acc[t], 0, 0, 0);
      if (ca) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ti, bs, acc[t], 0, 0, 0);
    }
  }

  /* acc_t[4q+i] = column 32(32t + r) + 8q + 4h + i: product dword w = 8(32t + r) + 2q + h
   * gets V_w = sum_i acc_t[4q+i] 2^(8i) (signed, |V_w| < 2^50), by signed 64-bit mads */
  int64_t *V = reinterpret_cast<int64_t *>(ws);
  int w8 = 1 << 8, w16 = 1 << 16, w24 = 1 << 24;
  asm volatile("" : "+s"(w8), "+s"(w16), "+s"(w24));   /* multiplies, not 64-bit shifts */
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = 32 * t + r;
    if (col >= 2 * NCH) continue;      /* past the product (partial last tile) */
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int64_t vq = (int64_t)acc[t][4 * q];
      vq = (int64_t)acc[t][4 * q + 1] * (int64_t)w8 + vq;
      vq = (int64_t)acc[t][4 * q + 2] * (int64_t)w16 + vq;
      vq = (int64_t)acc[t][4 * q + 3] * (int64_t)w24 + vq;
      const int w = 8 * col + 2 * q + h;
      V[w + (w >> 4)] = vq;            /* element w at w + w/16: conflict-free writes and reads */
    }
  }
  /* lane L: product limbs KP*L .. = low word of V_w plus the high part of V_(w-1) */
  const int w0 = KP * lane;
  int32_t hprev = 0;
  if (lane) hprev = reinterpret_cast<const int32_t *>(V)[2 * (w0 - 1 + ((w0 - 1) >> 4)) + 1];
  uint32_t D[KP];
  int64_t c = 0;
#pragma unroll
  for (int t = 0; t < KP; t++) {
    const int w = w0 + t;
    const int64_t vt = V[w + (w >> 4)];
    const int64_t sum = (int64_t)(uint32_t)vt + (int64_t)hprev + c;
    D[t] = (uint32_t)sum;
    c = sum >> 32;                     /* in [-1, 1] */
    hprev = (int32_t)(vt >> 32);
  }
  /* signed carry from the lane below, then the 1-bit carries/borrows it leaves; the carry
   * out of the top limb is the X^2 term and is dropped */
  const int ci = (int)g.from_prev((uint32_t)(int32_t)c);
  g.add_lane_words(D, ci > 0 ? (uint32_t)ci : 0u);
  g.sub_lane_words(D, ci < 0 ? (uint32_t)(-ci) : 0u);

  /* P in [2^(64n-2), 2^(64n)) for normal operands: if the top bit is clear, shift left once */
  const bool full = (__builtin_amdgcn_readlane(D[KP - 1], 63) >> 31) != 0;
  if (!full) {
    const uint32_t pv = g.from_prev(D[KP - 1]);
#pragma unroll
    for (int t = KP - 1; t > 0; t--) D[t] = __builtin_amdgcn_alignbit(D[t], D[t - 1], 31);
    D[0] = __builtin_amdgcn_alignbit(D[0], pv, 31);
  }
  /* mantissa = limbs n..2n-1 (lanes 32..63), round word = limb n-1, sticky = limbs below */
  const uint32_t rw = __builtin_amdgcn_readlane(D[KP - 1], 31);
  uint32_t below = 0;
#pragma unroll
  for (int t = 0; t < KP - 1; t++) below |= D[t];
  const uint32_t loc = lane < 31 ? (below | D[KP - 1]) : lane == 31 ? below : 0u;
  const uint32_t sticky = __ballot(loc != 0) != 0 ? 1u : 0u;
  const uint32_t sign = (ea ^ eb) & 1u;
  int64_t E = (int64_t)(ea >> 1) + (int64_t)(eb >> 1) - FPA_BIAS_LL - (full ? 0 : 1);
  if (round_mantissa<KP, 32>(half, D, (int)a.mode, sign, rw, sticky)) E += 1;
  store_result(a, s, half.gl, pack_exp(E, sign, 5, 0x7FFFFFFFll), D);
}

}  // namespace fpa

/* mul on the matrix cores for n = precision/32 limbs a multiple of 64; -1 otherwise */
extern "C" int fpa_impl_launch_mul_mfma(const fpa_arguments *a, void *stream) {
  if (a->count == 0) return 0;
  const dim3 grid((a->count + 3u) / 4u), block(256);
  hipStream_t st = (hipStream_t)stream;
  switch (a->precision) {
    case 2048u: hipLaunchKernelGGL(fpa::fpa_mul_mfma_kernel<8>, grid, block, 0, st, *a); break;
    case 4096u: hipLaunchKernelGGL(fpa::fpa_mul_mfma_kernel<16>, grid, block, 0, st, *a); break;
    case 6144u: hipLaunchKernelGGL(fpa::fpa_mul_mfma_kernel<24>, grid, block, 0, st, *a); break;
    case 8192u: hipLaunchKernelGGL(fpa::fpa_mul_mfma_kernel<32>, grid, block, 0, st, *a); break;
    case 16384u: hipLaunchKernelGGL(fpa::fpa_mul_mfma_kernel<64>, grid, block, 0, st, *a); break;
    default: return -1;
  }
  return (int)hipGetLastError();
}
