/*
 * Kernel wrapper shared by the per-op compilation units: one wavefront per 64/L
 * values, 4 waves per 256-thread block, each wave with its own LDS slice (no
 * workgroup barriers).  Value index v = wave * (64/L) + lane / L; out-of-range groups
 * read a valid element and store nothing.
 */
#ifndef FPA_KERNEL_COMMON_H
#define FPA_KERNEL_COMMON_H

#include "fpa_ops.h"
#include "fpa_shapes.h"

/* Occupancy target (waves per SIMD) of an element-wise kernel.  sqrt with 16 limbs on 8
 * lanes needs 130-140 VGPRs left alone (3 waves per SIMD); held to 128 it spills a little
 * outside the digit loop and runs faster at 4.  div with 32 limbs on 4 lanes needs about
 * 172 VGPRs; held to 168 (3 waves per SIMD) it spills a few words outside the loop and runs
 * 3-6% faster (with 8 or 16 lanes the same target spilled inside the loop and ran 11x
 * slower).  On 2 lanes the same target gave wrong low words in a few percent of the values
 * of large launches, differently from run to run (and none in small ones), so that shape
 * takes the compiler's own choice; the 4-lane one gives none on the same launches (a GPU
 * test runs every div and sqrt shape on 2^16 values).  Others take the compiler's own
 * choice (1 = no target).  A build can set one target for every kernel (make
 * EXTRA_HIPFLAGS=-DFPA_WAVES_PER_EU=n). */
__host__ __device__ constexpr int fpa_waves_per_eu(int op, int K, int L) {
#ifdef FPA_WAVES_PER_EU
  return (void)op, (void)K, (void)L, FPA_WAVES_PER_EU;
#else
  return op == 5 && K == 16 && L == 8 ? 4 : op == 4 && K == 32 && L == 4 ? 3 : 1;
#endif
}
#define FPA_KERNEL_ATTRS(OP, K, L) __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fpa_waves_per_eu(OP, K, L))))

namespace fpa {

/* LDS words of scratch per value for each op (neg, add, sub, mul, div, sqrt) */
template <int OP, int K, int L>
struct ScratchWords {
  static constexpr int n = K * L;
  static constexpr int raw = OP == 0 ? 0
                           : OP <= 2 ? n + 1
                           : OP == 3 ? mul_op_lds_words<K, L>()
                           : OP == 4 ? div_lds_words(K, L)
                           : sqrt_lazy_lds_words<K, L>();
  /* pad to a multiple of 4 words, plus 4 so successive values start on different banks.
   * div and sqrt: an even stride of 2 mod 4 words instead, so the 64/L values of a wave
   * start on 64/L different banks of the 32 (their digit steps and relayouts have every
   * group access the same offset at once; a stride of 0 mod 32, which the rule above gives
   * for some shapes, put all of them on one bank) */
  static constexpr int quad = ((raw + 3) / 4) * 4 + 4;
  static constexpr int even = ((raw + 1) / 2) * 2;
  static constexpr int value = OP >= 4 ? (even % 4 == 2 ? even : even + 2) : quad;
};

/* Waves per block: 4, or 1 for mul with a short product, whose LDS then packs by the
 * wave (10 waves per CU at 4096 bits instead of 2 blocks of 4). */
template <int OP, int K, int L>
__host__ __device__ constexpr int fpa_waves_per_block() { return OP == 3 && mul_short_ok<K, L>() ? 1 : 4; }

template <int K, int L, int OP>
__global__ FPA_KERNEL_ATTRS(OP, K, L) void fpa_kernel(fpa_arguments a) {
  constexpr int VPW = 64 / L;
  constexpr int WPB = fpa_waves_per_block<OP, K, L>();
  constexpr int SW = ScratchWords<OP, K, L>::value;
  __shared__ __attribute__((aligned(16))) uint32_t smem[WPB * VPW * SW];
  Group<L> g;
  const int wib = threadIdx.x >> 6;
  const int gid = g.lane / L;
  const uint64_t vv = ((uint64_t)blockIdx.x * WPB + wib) * VPW + gid;
  Slot s;
  s.valid = vv < a.count;
  s.v = s.valid ? (uint32_t)vv : a.count - 1;
  uint32_t *lds = smem + (wib * VPW + gid) * SW;
  if constexpr (OP == 0) op_neg<K, L>(a, s, g);
  else if constexpr (OP == 1) op_addsub<K, L>(a, s, g, lds, false);
  else if constexpr (OP == 2) op_addsub<K, L>(a, s, g, lds, true);
  else if constexpr (OP == 3) op_mul<K, L>(a, s, g, lds, SW);
  else if constexpr (OP == 4) op_div<K, L>(a, s, g, lds);
  else op_sqrt<K, L>(a, s, g, lds);
}

template <int K, int L, int OP>
static int launch_shape(const fpa_arguments &a, hipStream_t st) {
  constexpr int VPW = 64 / L;
  if (a.count == 0) return 0;
  constexpr int WPB = fpa_waves_per_block<OP, K, L>();
  uint64_t waves = ((uint64_t)a.count + VPW - 1) / VPW;
  uint64_t blocks = (waves + WPB - 1) / WPB;
  hipLaunchKernelGGL((fpa_kernel<K, L, OP>), dim3((uint32_t)blocks), dim3(64 * WPB), 0, st, a);
  return (int)hipGetLastError();
}

}  // namespace fpa

namespace fpa {

/* The compiled shapes as data, in FPA_SHAPES order. */
struct ShapeNKL { int n, K, L; };
#define FPA_SHAPE_NKL(n_, K_, L_) ShapeNKL{n_, K_, L_},
constexpr ShapeNKL kShapes[] = { FPA_SHAPES(FPA_SHAPE_NKL) };
#undef FPA_SHAPE_NKL
constexpr int kNumShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

/* Shapes an op is compiled for: mul stops at FPA_MUL_MAX_K limbs per lane. */
constexpr bool op_has_shape(int op, int K) { return op != 3 || K <= FPA_MUL_MAX_K; }

/* Compilation parts.  An op's shapes can be split over several files that build in
 * parallel; shapes are dealt to parts heaviest first, each to the least-loaded part, by a
 * rough compile-time weight (div and sqrt with 24-32 limbs per lane: 10-20 s per shape;
 * with 16: about 8 s; most others 4-6 s). */
constexpr int compile_weight(int op, int K) {
  return op >= 4 ? (K >= 24 ? 3 : K >= 16 ? 2 : 1) : 1;
}
constexpr int compile_part(int op, int K, int L, int nparts) {
  int load[16] = {};
  int part = -1;
  for (int w = 3; w >= 1; w--) {
    for (int i = 0; i < kNumShapes; i++) {
      if (!op_has_shape(op, kShapes[i].K) || compile_weight(op, kShapes[i].K) != w) continue;
      int best = 0;
      for (int q = 1; q < nparts; q++)
        if (load[q] < load[best]) best = q;
      load[best] += w;
      if (kShapes[i].K == K && kShapes[i].L == L) part = best;
    }
  }
  return part;
}

/* Launch (K, L) if this part compiles it; -1 if it does not.  A template, so the shapes
 * of other parts are never instantiated here. */
template <int OP, int PART, int NPARTS>
static int launch_part(const fpa_arguments *a, int K, int L, hipStream_t st) {
#define FPA_LAUNCH_CASE_PART(n, K_, L_)                                   \
  if constexpr (op_has_shape(OP, K_) && compile_part(OP, K_, L_, NPARTS) == PART) \
    if (K == K_ && L == L_) return launch_shape<K_, L_, OP>(*a, st);
  FPA_SHAPES(FPA_LAUNCH_CASE_PART)
#undef FPA_LAUNCH_CASE_PART
  return -1;
}

}  // namespace fpa

/* defines extern "C" int fpa_impl_launch_<name>_part<PART>(const fpa_arguments*, int K,
 * int L, void *stream): launches the shapes of compilation part PART of NPARTS, returns
 * -1 for a shape of another part */
#define FPA_DEFINE_LAUNCHER_PART(name, OPV, PART, NPARTS)                                            \
  extern "C" int fpa_impl_launch_##name##_part##PART(const fpa_arguments *a, int K, int L, void *stream) { \
    return fpa::launch_part<OPV, PART, NPARTS>(a, K, L, (hipStream_t)stream);                           \
  }

#endif
