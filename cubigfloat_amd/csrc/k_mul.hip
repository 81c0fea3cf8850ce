/* Kernel instantiations for fpa_mul (all its compiled shapes). */
#include "fpa_kernel_common.h"

FPA_DEFINE_LAUNCHER_PART(mul, 3, 0, 1)

/* Product rounds per lane of shape (K, L): R1 + R2 of its lane schedule for a short
 * product, L for the full product scan; 0 for a shape mul is not compiled for.  The batch
 * planner weighs segments with it. */
extern "C" int fpa_impl_mul_rounds(int K, int L) {
#define FPA_MUL_ROUNDS_CASE(n, K_, L_)                                                       \
  if constexpr (fpa::op_has_shape(3, K_))                                                  \
    if (K == K_ && L == L_)                                                                \
      return fpa::mul_rounds<K_, L_>();
  FPA_SHAPES(FPA_MUL_ROUNDS_CASE)
#undef FPA_MUL_ROUNDS_CASE
  return 0;
}
