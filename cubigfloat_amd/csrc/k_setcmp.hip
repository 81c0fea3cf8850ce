/*
 * fpa_sgn, fpa_cmp and fpa_set (fpa.h:57-61).
 *
 * sgn: one lane per value, a class-table lookup (fpa_threaded.h:1-12, 26-43).
 * cmp: one group of L lanes per value pair.  The class table (fpa_threaded.h:14-24)
 *      decides unless both operands are values of the same sign; then the exponent
 *      words decide, and equal words fall to a group-wide mantissa compare: each lane
 *      finds the highest differing limb of its chunk, and the highest lane that found
 *      one gives the answer (fpa_threaded.h:45-93 does the same per warp).
 * set: one group per destination value (shape of the destination precision).  The
 *      destination's n limbs are the source's top n limbs (zeros below a shorter
 *      source); the source limb just below them is the round word and the rest OR into
 *      the sticky bit, then the usual rounding with the add/sub/mul range rule.  The
 *      reference's set kernel is unfinished (fpa_warp_set.h:134-199); results equal
 *      mpfr_set (its gold, tests.c:580-605).  Special classes are copied.
 * set_ui / set_si / set_float / set_double: one group per destination value; every lane
 *      of the group reads the value's 4- or 8-byte source, fpa_encode_scalar gives the
 *      exponent word and the left-aligned significand, the group's top lane writes it as
 *      its two top limbs and every other limb is zero (reference fpa_warp_set.h:1-132:
 *      one warp per value, the same words).
 */
#include "fpa_kernel_common.h"

namespace fpa {

/* entries -1, 0, +1; 2 = compare the values; 3 = error (result 0) */
__constant__ int8_t k_sgn_tab[9] = { 1, -1, 0, -1, 1, -1, 1, 3, 3 };
__constant__ int8_t k_cmp_tab[81] = {
  2, 1, 1, 1, 1, 1, -1, 3, 3,
  -1, 2, -1, -1, -1, 1, -1, 3, 3,
  -1, 1, 0, 1, -1, 1, -1, 3, 3,
  -1, 1, -1, 3, -1, 1, -1, 3, 3,
  -1, 1, 1, 1, 3, 1, -1, 3, 3,
  -1, -1, -1, -1, -1, 3, -1, 3, 3,
  1, 1, 1, 1, 1, 1, 3, 3, 3,
  3, 3, 3, 3, 3, 3, 3, 3, 3,
  3, 3, 3, 3, 3, 3, 3, 3, 3,
};

__global__ __launch_bounds__(256) void fpa_sgn_kernel(uint32_t count, const uint32_t *x_exps, int32_t *r) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const int t = k_sgn_tab[fpa_classify(x_exps[i])];
  r[i] = t == 3 ? 0 : t;
}

/* value slot of this lane's group (same mapping as the element-wise kernels) */
template <int L>
__device__ __forceinline__ Slot group_slot(const Group<L> &g, uint32_t count) {
  constexpr int VPW = 64 / L;
  const int wib = threadIdx.x >> 6;
  const uint64_t vv = ((uint64_t)blockIdx.x * 4 + wib) * VPW + g.lane / L;
  Slot s;
  s.valid = vv < count;
  s.v = s.valid ? (uint32_t)vv : count - 1;
  return s;
}

template <int K, int L>
__global__ __launch_bounds__(256) void fpa_cmp_kernel(fpa_arguments a) {
  Group<L> g;
  const Slot s = group_slot<L>(g, a.count);
  const uint32_t ea = a.a_exps[s.v], eb = a.b_exps[s.v];
  uint32_t A[K], B[K];
  load_value<K, L>(g, a.a_limbs, s, A);
  load_value<K, L>(g, a.b_limbs, s, B);
  /* this lane's chunk: +1 / -1 at its highest differing limb, 0 if equal */
  int c = 0;
#pragma unroll
  for (int i = K - 1; i >= 0; i--)
    if (c == 0 && A[i] != B[i]) c = A[i] > B[i] ? 1 : -1;
  const uint64_t diff = g.ballot(c != 0);
  const int mant = diff ? (int)g.bcast((uint32_t)c, 63 - __clzll(diff)) : 0;
  int t = k_cmp_tab[fpa_classify(ea) * 9 + fpa_classify(eb)];
  if (t == 2) {
    const int sgn = (ea & 1u) ? -1 : 1;
    t = ea != eb ? (ea > eb ? sgn : -sgn) : sgn * mant;
  }
  if (s.valid && g.gl == 0) reinterpret_cast<int32_t *>(a.r_exps)[s.v] = t == 3 ? 0 : t;
}

template <int K, int L>
__global__ __launch_bounds__(256) void fpa_set_kernel(fpa_arguments a) {
  constexpr int n = K * L;
  Group<L> g;
  const Slot s = group_slot<L>(g, a.count);
  const int nx = (int)(a.set_source_precision >> 5);
  const int off = nx - n;                        /* source limb of destination limb 0 */
  const uint32_t *x = a.a_limbs + (size_t)s.v * nx;
  const uint32_t ex = a.a_exps[s.v];
  uint32_t H[K];
#pragma unroll
  for (int i = 0; i < K; i++) {
    const int src = g.gl * K + i + off;
    H[i] = src >= 0 ? x[src] : 0u;
  }
  const uint32_t rw = off >= 1 ? x[off - 1] : 0u;
  uint32_t loc = 0;
  for (int w = g.gl; w < off - 1; w += L) loc |= x[w];
  const uint32_t sticky = g.any(loc != 0) ? 1u : 0u;
  uint32_t re;
  if (ex <= FPA_C_SNAN) {
    re = ex;
  } else {
    const uint32_t sign = ex & 1u;
    int64_t E = (int64_t)(ex >> 1);
    if (round_mantissa<K, L>(g, H, (int)a.mode, sign, rw, sticky)) E += 1;
    re = pack_exp(E, sign, 5, 0x7FFFFFFFll);
  }
  store_result(a, s, g.gl, re, H);
}

template <int K, int L>
__global__ __launch_bounds__(256) void fpa_set_scalar_kernel(fpa_arguments a, int kind) {
  static_assert(K >= 2, "the top lane holds the significand's two words");
  Group<L> g;
  const Slot s = group_slot<L>(g, a.count);
  const uint64_t bits = kind == FPA_SRC_F32 ? (uint64_t)static_cast<const uint32_t *>(a.values)[s.v]
                                            : static_cast<const uint64_t *>(a.values)[s.v];
  uint64_t top;
  const uint32_t re = fpa_encode_scalar(kind, bits, &top);
  uint32_t x[K];
#pragma unroll
  for (int i = 0; i < K; i++) x[i] = 0;
  if (g.gl == L - 1) {
    x[K - 1] = (uint32_t)(top >> 32);
    x[K - 2] = (uint32_t)top;
  }
  store_result(a, s, g.gl, re, x);
}

}  // namespace fpa

extern "C" int fpa_impl_launch_sgn(uint32_t count, const uint32_t *x_exps, int32_t *r, void *stream) {
  if (count == 0) return 0;
  hipLaunchKernelGGL(fpa::fpa_sgn_kernel, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, count,
                     x_exps, r);
  return (int)hipGetLastError();
}

/* over args->count values with shape (K, L) of args->precision: cmp (op 0), set (op 1), or
 * set from the 4/8-byte sources at args->values (op 2 + FPA_SRC_*: ui, si, float, double) */
extern "C" int fpa_impl_launch_setcmp(int op, const fpa_arguments *args, int K, int L, void *stream) {
  if (args->count == 0) return 0;
  if (op < 0 || op > 2 + FPA_SRC_F64) return -1;
#define FPA_SETCMP_CASE(n_, K_, L_)                                                               \
  if (args->precision == 32u * n_ && K == K_ && L == L_) {                                        \
    uint64_t waves = ((uint64_t)args->count + (64 / L_) - 1) / (64 / L_);                         \
    dim3 grid((uint32_t)((waves + 3) / 4));                                                       \
    if (op == 0) hipLaunchKernelGGL((fpa::fpa_cmp_kernel<K_, L_>), grid, dim3(256), 0, (hipStream_t)stream, *args); \
    else if (op == 1) hipLaunchKernelGGL((fpa::fpa_set_kernel<K_, L_>), grid, dim3(256), 0, (hipStream_t)stream, *args); \
    else hipLaunchKernelGGL((fpa::fpa_set_scalar_kernel<K_, L_>), grid, dim3(256), 0, (hipStream_t)stream, *args, \
                            op - 2);                                                              \
    return (int)hipGetLastError();                                                                \
  }
  FPA_SHAPES(FPA_SETCMP_CASE)
#undef FPA_SETCMP_CASE
  return -1;
}
