/* Kernel instantiations for fpa_mul (all compiled shapes). */
#include "fpa_kernel_common.h"

#define FPA_LAUNCH_CASE_mul(n, K_, L_) \
  if (K == K_ && L == L_) return fpa::launch_shape<K_, L_, 3>(*a, st);
FPA_DEFINE_LAUNCHER(mul, 3)
