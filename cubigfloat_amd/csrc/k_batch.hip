/*
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
    /* A group of 16 lanes is one DPP row: the lockstep window shift reads no other group,
     * so it stays exact while the wave's groups run chains of different lengths. */
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
