/* Kernel instantiations for fpa_sqrt (all compiled shapes). */
#include "fpa_kernel_common.h"

#define FPA_LAUNCH_CASE_sqrt(n, K_, L_) \
  if (K == K_ && L == L_) return fpa::launch_shape<K_, L_, 5>(*a, st);
FPA_DEFINE_LAUNCHER(sqrt, 5)
