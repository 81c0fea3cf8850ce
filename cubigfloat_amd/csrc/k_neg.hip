/* Kernel instantiations for fpa_neg (all compiled shapes). */
#include "fpa_kernel_common.h"

#define FPA_LAUNCH_CASE_neg(n, K_, L_) \
  if (K == K_ && L == L_) return fpa::launch_shape<K_, L_, 0>(*a, st);
FPA_DEFINE_LAUNCHER(neg, 0)
