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
#include <mutex>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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

/* ------------------ fpa_set_* from host values without allocating ------------------ */

/* Staging for fpa_set_scalars_noalloc: the 4- or 8-byte sources of one call.  A file-scope
 * device array: each GPU the library runs on holds its own copy (8 B x
 * FPA_NOALLOC_MAX_VALUES = 4 MiB) until the process ends.  Host code never passes the
 * array's name to a kernel or a copy (on the host it is only a stand-in); it uses the
 * device address that hipGetSymbolAddress returns for the current device. */
static __device__ __align__(256) uint64_t scratch[FPA_NOALLOC_MAX_VALUES];

#define FPA_NOALLOC_MAX_DEVICES 64
static void *scratch_ptr[FPA_NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

/* device address of `scratch` on the current device, looked up once per device */
static void *scratch_on_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= FPA_NOALLOC_MAX_DEVICES) {
    fprintf(stderr, "\nfpa_set_scalars_noalloc: no usable current device (%d, %s)\n", dev, hipGetErrorString(e));
    exit(1);
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *p = NULL;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "\nfpa_set_scalars_noalloc: hipGetSymbolAddress failed on device %d: %s\n", dev,
              hipGetErrorString(e));
      exit(1);
    }
    scratch_ptr[dev] = p;
  }
  return scratch_ptr[dev];
}

extern "C" fpa_result fpa_set_scalars_noalloc(fpa_t r, int kind, const void *values) {
  if (kind < FPA_SRC_U64 || kind > FPA_SRC_F64) return fpa_math_error;
  if (!fpa_precision_supported(r->precision)) return fpa_invalid_precisions;
  if (r->count > FPA_NOALLOC_MAX_VALUES) {
    fprintf(stderr, "fpa_set_scalars_noalloc: %u values exceed the staging limit of %u values\n", r->count,
            (unsigned)FPA_NOALLOC_MAX_VALUES);
    return fpa_invalid_counts;
  }
  if (r->count == 0) return fpa_success;
  void *ptr = scratch_on_device();
  hipStream_t st = (hipStream_t)fpa_get_stream();
  const size_t bytes = (kind == FPA_SRC_F32 ? 4u : 8u) * (size_t)r->count;
  /* every staged word the kernel reads is written here first: nothing needs clearing */
  hipError_t e = hipMemcpyAsync(ptr, values, bytes, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    fprintf(stderr, "\nfpa_set_scalars_noalloc: staging copy failed: %s\n", hipGetErrorString(e));
    exit(1);
  }
  fpa_arguments a;
  memset(&a, 0, sizeof a);
  a.count = r->count;
  a.precision = r->precision;
  a.r_exps = r->gpu_exps;
  a.r_limbs = r->gpu_limbs;
  a.values = ptr;
  int K = 0, L = 0;
  fpa_kernel_variant_shape(FPA_OP_NEG, r->precision,
                           fpa_kernel_get_variant_for_count(FPA_OP_NEG, r->precision, r->count), &K, &L);
  const int rc = fpa_impl_launch_setcmp(2 + kind, &a, K, L, st);
  if (rc != 0) {
    fprintf(stderr, "\nfpa_set_scalars_noalloc: launch failed (%d)\n", rc);
    exit(1);
  }
  /* the next call on this device overwrites the staging words: wait for the kernel */
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    fprintf(stderr, "\nfpa_set_scalars_noalloc: %s\n", hipGetErrorString(e));
    exit(1);
  }
  return fpa_success;
}
