/*
 * Batched execution of mixed-precision element-wise ops, and fused dot-product chains.
 *
 * fpa_batch_kernel is persistent: a fixed grid sized to the device's resident capacity;
 * each wavefront walks a host-built list of tiles (tile = the 64/L values one wave
 * handles for one (op, precision) segment), striding by the number of waves in the grid.
 * The tile list is sorted by estimated cost, so the expensive div/sqrt/high-precision
 * tiles start first on every wave.  The per-tile branch is on a wave-uniform kind, so
 * each wave runs one specialised (K, L) code path at a time.  No counters or atomics:
 * the launch can be captured in a hipGraph and replayed without resetting anything.
 *
 * fpa_dot_kernel: one group of L lanes per dot product, accumulator kept in registers;
 * acc_{i+1} = RN(acc_i + RN(x_i * y_i)) in index order.  Groups take dot products in
 * the host-given order (longest first) so the groups sharing a wave have similar lengths.
 */
#include "fpa_kernel_common.h"
#include "fpa_batch_shapes.h"
#include "../../include/fpa_batch.h"

namespace fpa {

/* dense kind index for every X(op, n, K, L) entry */
enum BatchKind {
#define FPA_KIND_ENUM(op, n, K, L) kind_##op##_##n,
  FPA_BATCH_SHAPES(FPA_KIND_ENUM)
#undef FPA_KIND_ENUM
  kind_count
};

/* LDS words one wavefront needs for a kind (64/L values) */
template <int OP, int K, int L>
struct WaveLds {
  static constexpr int value = (64 / L) * ScratchWords<OP, K, L>::value;
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int batch_wave_lds() {
  int m = 4;
#define FPA_KIND_LDS(op, n, K, L) m = cmax(m, WaveLds<op, K, L>::value);
  FPA_BATCH_SHAPES(FPA_KIND_LDS)
#undef FPA_KIND_LDS
  return m;
}
constexpr int kBatchWaveLds = batch_wave_lds();

template <int K, int L, int OP>
__device__ __forceinline__ void run_tile(const fpa_arguments &a, uint32_t first, uint32_t *wave_lds) {
  constexpr int SW = ScratchWords<OP, K, L>::value;
  Group<L> g;
  const int gid = g.lane / L;
  const uint64_t vv = (uint64_t)first + gid;
  Slot s;
  s.valid = vv < a.count;
  s.v = s.valid ? (uint32_t)vv : a.count - 1;
  uint32_t *lds = wave_lds + gid * SW;
  if constexpr (OP == 1) op_addsub<K, L>(a, s, g, lds, false);
  else if constexpr (OP == 2) op_addsub<K, L>(a, s, g, lds, true);
  else if constexpr (OP == 3) op_mul<K, L>(a, s, g, lds);
  else if constexpr (OP == 4) op_div<K, L>(a, s, g, lds);
  else op_sqrt<K, L>(a, s, g, lds);
}

__global__ __launch_bounds__(256) void fpa_batch_kernel(const fpa_batch_segment *segs, const uint2 *tiles,
                                                        uint32_t ntiles) {
  __shared__ __attribute__((aligned(16))) uint32_t smem[4 * kBatchWaveLds];
  const int wib = threadIdx.x >> 6;
  uint32_t *wave_lds = smem + wib * kBatchWaveLds;
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t t = blockIdx.x * 4 + wib; t < ntiles; t += nwaves) {
    const uint2 tile = tiles[t];
    const fpa_batch_segment &sg = segs[tile.x];
    const fpa_arguments a = sg.args;
    switch (sg.kind) {
#define FPA_KIND_CASE(op, n, K, L) \
      case kind_##op##_##n: run_tile<K, L, op>(a, tile.y, wave_lds); break;
      FPA_BATCH_SHAPES(FPA_KIND_CASE)
#undef FPA_KIND_CASE
      default: break;
    }
  }
}

/* ------------------------------ dot chains ------------------------------ */
template <int K, int L>
__global__ __launch_bounds__(256) void fpa_dot_kernel(fpa_dot_arguments d) {
  constexpr int VPW = 64 / L;
  constexpr int SW = cmax(ScratchWords<3, K, L>::value, ScratchWords<1, K, L>::value);
  __shared__ __attribute__((aligned(16))) uint32_t smem[4 * VPW * SW];
  Group<L> g;
  const int wib = threadIdx.x >> 6;
  const int gid = g.lane / L;
  const uint64_t slot = ((uint64_t)blockIdx.x * 4 + wib) * VPW + gid;
  if (slot >= d.count) return;               /* whole groups leave together */
  uint32_t *lds = smem + (wib * VPW + gid) * SW;
  const uint32_t j = d.order[slot];
  const uint64_t k0 = d.offsets[j], k1 = d.offsets[j + 1];
  constexpr int n = K * L;
  uint32_t acc_e = FPA_C_ZERO, acc[K];
#pragma unroll
  for (int i = 0; i < K; i++) acc[i] = 0;
  for (uint64_t k = k0; k < k1; k++) {
    uint32_t X[K], Y[K], P[K], pe;
    load_chunk(X, d.x_limbs + k * n + (size_t)g.gl * K);
    load_chunk(Y, d.y_limbs + k * n + (size_t)g.gl * K);
    core_mul<K, L>(g, lds, d.mode, d.x_exps[k], X, d.y_exps[k], Y, pe, P);
    uint32_t se, S[K];
    core_addsub<K, L>(g, lds, d.mode, false, acc_e, acc, pe, P, se, S);
    acc_e = se;
#pragma unroll
    for (int i = 0; i < K; i++) acc[i] = S[i];
  }
  uint32_t z[K];
#pragma unroll
  for (int i = 0; i < K; i++) z[i] = acc_e > FPA_C_SNAN ? acc[i] : 0u;
  store_chunk(d.r_limbs + (size_t)j * n + (size_t)g.gl * K, z);
  if (g.gl == 0) d.r_exps[j] = acc_e;
}

}  // namespace fpa

extern "C" int fpa_impl_batch_kind(int op, uint32_t precision) {
#define FPA_KIND_FIND(op_, n, K, L) \
  if (op == op_ && precision == 32u * n) return fpa::kind_##op_##_##n;
  FPA_BATCH_SHAPES(FPA_KIND_FIND)
#undef FPA_KIND_FIND
  return -1;
}

extern "C" int fpa_impl_batch_kind_shape(int kind, int *op, uint32_t *precision, int *K, int *L) {
  int i = 0;
#define FPA_KIND_SHAPE(op_, n, K_, L_)                                             \
  if (i++ == kind) { *op = op_; *precision = 32u * n; *K = K_; *L = L_; return 0; }
  FPA_BATCH_SHAPES(FPA_KIND_SHAPE)
#undef FPA_KIND_SHAPE
  return -1;
}

/* resident 256-thread blocks per CU for the persistent kernel */
extern "C" int fpa_impl_batch_blocks_per_cu(void) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fpa::fpa_batch_kernel, 256, 0) != hipSuccess) return 1;
  return blocks > 0 ? blocks : 1;
}

extern "C" int fpa_impl_batch_launch(const fpa_batch_segment *segs, const void *tiles, uint32_t ntiles,
                                     uint32_t grid, void *stream) {
  if (ntiles == 0) return 0;
  hipLaunchKernelGGL(fpa::fpa_batch_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, segs,
                     (const uint2 *)tiles, ntiles);
  return (int)hipGetLastError();
}

extern "C" int fpa_impl_dot_launch(const fpa_dot_arguments *d, void *stream) {
  if (d->count == 0) return 0;
#define FPA_DOT_CASE(op_, n, K, L)                                                               \
  if (op_ == 3 && d->precision == 32u * n) {                                                     \
    uint64_t waves = ((uint64_t)d->count + (64 / L) - 1) / (64 / L);                             \
    hipLaunchKernelGGL((fpa::fpa_dot_kernel<K, L>), dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, \
                       (hipStream_t)stream, *d);                                                 \
    return (int)hipGetLastError();                                                               \
  }
  FPA_BATCH_SHAPES(FPA_DOT_CASE)
#undef FPA_DOT_CASE
  return -1;
}
