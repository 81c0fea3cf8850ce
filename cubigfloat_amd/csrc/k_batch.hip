/*
 * Persistent mixed batch kernel and fused dot-product chains.
 *
 * fpa_batch_persistent_kernel: one launch for a whole mixed batch.  A fixed grid of
 * resident blocks; each wavefront takes tiles from a work queue (one atomic add per tile)
 * until the queue is empty.  A tile is one wavefront's 64/L values of one (op,
 * precision, mode) segment, run with the segment's kind -- an (op, K, L) shape from
 * FPA_PERSIST_SHAPES, every one inlined into this kernel.  The host lists the tiles most
 * expensive first, so the long tiles (high-precision div and sqrt) start at once on every
 * CU and the short ones fill in behind them.  The queue resets itself: the last wave to
 * finish (counted in the queue's second word) sets both words back to zero, so the launch
 * can be captured in a hipGraph and replayed with no memset or reset node.  The segment
 * table (at most FPA_PERSIST_MAX_SEGMENTS segments) is a kernel argument.
 *
 * Fused dot-product chains.
 *
 * fpa_dot_kernel: one group of L lanes per dot product, accumulator kept in registers;
 * acc_{i+1} = RN(acc_i + RN(x_i * y_i)) in index order.  Groups take dot products in
 * the host-given order (longest first) so the groups sharing a wave have similar lengths.
 * Instantiated for the compiled shapes with K <= FPA_DOT_MAX_K; the host picks the
 * dispatcher's FPA_KERNEL_DOT shape.
 */
#include "fpa_kernel_common.h"
#include "../../include/fpa_batch.h"

namespace fpa {

constexpr int cmax(int a, int b) { return a > b ? a : b; }

/* ----------------------------- persistent batch ----------------------------- */
/* kind index for every X(op, n, K, L) entry */
enum PersistKind {
#define FPA_PKIND_ENUM(op, n, K, L) pkind_##op##_##n,
  FPA_PERSIST_SHAPES(FPA_PKIND_ENUM)
#undef FPA_PKIND_ENUM
  pkind_count
};

/* LDS words one wavefront needs for a kind (64/L values) */
constexpr int persist_wave_lds() {
  int m = 4;
#define FPA_PKIND_LDS(op, n, K, L) m = cmax(m, (64 / L) * ScratchWords<op, K, L>::value);
  FPA_PERSIST_SHAPES(FPA_PKIND_LDS)
#undef FPA_PKIND_LDS
  return m;
}
constexpr int kPersistWaveLds = persist_wave_lds();

template <int OP, int K, int L>
__device__ __forceinline__ void run_tile(const fpa_arguments &a, uint32_t first, uint32_t *wave_lds) {
  constexpr int SW = ScratchWords<OP, K, L>::value;
  Group<L> g;
  /* the lane number passes through an empty volatile asm, so nothing derived from it is
   * hoisted out of the tile loops to the kernel entry: that would keep every kind's lane
   * constants live across the whole kernel (measured: 179 VGPRs instead of 63 for the six
   * add kinds together) */
  {
    uint32_t lane = (uint32_t)g.lane;
    asm volatile("" : "+v"(lane));
    g.lane = (int)lane;
    g.gl = g.lane & (L - 1);
    g.gbase = g.lane - g.gl;
  }
  const int gid = g.lane / L;
  const uint64_t vv = (uint64_t)first + gid;
  Slot s;
  s.valid = vv < a.count;
  s.v = s.valid ? (uint32_t)vv : a.count - 1;
  uint32_t *lds = wave_lds + gid * SW;
  if constexpr (OP == 1) op_addsub<K, L>(a, s, g, lds, false);
  else if constexpr (OP == 2) op_addsub<K, L>(a, s, g, lds, true);
  else if constexpr (OP == 3) op_mul<K, L>(a, s, g, lds, SW);
  else if constexpr (OP == 4) op_div<K, L>(a, s, g, lds);
  else op_sqrt<K, L>(a, s, g, lds);
}

/* next tile index for this wave (lane 0 takes it from the queue) */
__device__ __forceinline__ uint32_t next_tile(uint32_t *queue) {
  uint32_t t = 0;
  if (__lane_id() == 0) t = atomicAdd(queue, 1u);
  return __builtin_amdgcn_readfirstlane(t);
}

__device__ __forceinline__ const uint2 *tiles_of(const fpa_persist_arguments &P) {
  return reinterpret_cast<const uint2 *>(P.tiles);
}

/* tiles of one kind, taken until the queue is empty or the next tile has another kind */
template <int OP, int K, int L>
__device__ __forceinline__ uint32_t run_kind(const fpa_persist_arguments &P, uint32_t t, uint32_t kind,
                                             uint32_t *wave_lds) {
  do {
    const uint2 tile = tiles_of(P)[t];
    run_tile<OP, K, L>(P.seg[__builtin_amdgcn_readfirstlane(tile.x)], __builtin_amdgcn_readfirstlane(tile.y),
                       wave_lds);
    t = next_tile(P.queue);
  } while (t < P.ntiles && P.kind[__builtin_amdgcn_readfirstlane(tiles_of(P)[t].x)] == kind);
  return t;
}

/* The segment table travels in the kernel arguments: pointers read from there are known to
 * be global, so every operand access is a global instruction (read from a table in memory
 * they would be generic pointers and FLAT accesses). */
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void fpa_batch_persistent_kernel(
    fpa_persist_arguments P) {
  __shared__ __attribute__((aligned(16))) uint32_t smem[4 * kPersistWaveLds];
  const int wib = threadIdx.x >> 6;
  uint32_t *wave_lds = smem + wib * kPersistWaveLds;
  uint32_t t = next_tile(P.queue);
  while (t < P.ntiles) {
    const uint32_t kind = P.kind[__builtin_amdgcn_readfirstlane(tiles_of(P)[t].x)];
    switch (kind) {
#define FPA_PKIND_CASE(op, n, K, L) \
      case pkind_##op##_##n: t = run_kind<op, K, L>(P, t, kind, wave_lds); break;
      FPA_PERSIST_SHAPES(FPA_PKIND_CASE)
#undef FPA_PKIND_CASE
      default: t = next_tile(P.queue); break;
    }
  }
  /* every wave has taken its last tile index; the last one to get here resets the queue */
  __threadfence();
  if (__lane_id() == 0) {
    const uint32_t done = atomicAdd(P.queue + 1, 1u);
    if (done == gridDim.x * 4u - 1u) {
      atomicExch(P.queue, 0u);
      atomicExch(P.queue + 1, 0u);
    }
  }
}

/* ------------------------------ dot chains ------------------------------ */
template <int K, int L>
__global__ __launch_bounds__(256) void fpa_dot_kernel(fpa_dot_arguments d) {
  constexpr int VPW = 64 / L;
  /* the mul core's scratch (with the short product's columns where the shape has one) and
   * the add core's */
  constexpr int SW = cmax(((mul_op_lds_words<K, L>() + 3) / 4) * 4 + 4, ScratchWords<1, K, L>::value);
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
    /* The short product where the shape has one (each group decides its own fallback);
     * else the full scan: a group of 16 lanes is one DPP row, whose lockstep window shift
     * reads no other group, so it stays exact while the wave's groups run chains of
     * different lengths. */
    if constexpr (mul_short_ok<K, L>())
      core_mul_short<K, L, false>(g, lds, SW, d.mode, d.x_exps[k], X, d.y_exps[k], Y, pe, P);
    else
      core_mul<K, L, L == 16>(g, lds, SW, d.mode, d.x_exps[k], X, d.y_exps[k], Y, pe, P);
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

/* kind of (op, precision) in the persistent kernel, -1 if it has none; its shape in K, L */
extern "C" int fpa_impl_persist_kind(int op, uint32_t precision, int *K, int *L) {
#define FPA_PKIND_FIND(op_, n, K_, L_)                                       \
  if (op == op_ && precision == 32u * n) { *K = K_; *L = L_; return fpa::pkind_##op_##_##n; }
  FPA_PERSIST_SHAPES(FPA_PKIND_FIND)
#undef FPA_PKIND_FIND
  return -1;
}

/* resident 256-thread blocks per CU of the persistent kernel */
extern "C" int fpa_impl_persist_blocks_per_cu(void) {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fpa::fpa_batch_persistent_kernel, 256, 0) != hipSuccess)
    return 1;
  return blocks > 0 ? blocks : 1;
}

extern "C" int fpa_impl_persist_launch(const fpa_persist_arguments *P, uint32_t grid, void *stream) {
  if (P->ntiles == 0) return 0;
  hipLaunchKernelGGL(fpa::fpa_batch_persistent_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, *P);
  return (int)hipGetLastError();
}

extern "C" int fpa_impl_dot_launch(const fpa_dot_arguments *d, int K, int L, void *stream) {
  if (d->count == 0) return 0;
#define FPA_DOT_CASE(n, K_, L_)                                                                  \
  if constexpr (K_ <= FPA_DOT_MAX_K)                                                             \
  if (d->precision == 32u * n && K == K_ && L == L_) {                                           \
    uint64_t waves = ((uint64_t)d->count + (64 / L_) - 1) / (64 / L_);                           \
    hipLaunchKernelGGL((fpa::fpa_dot_kernel<K_, L_>), dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, \
                       (hipStream_t)stream, *d);                                                 \
    return (int)hipGetLastError();                                                               \
  }
  FPA_SHAPES(FPA_DOT_CASE)
#undef FPA_DOT_CASE
  return -1;
}
