/* Kernel instantiations for fpa_sub (all compiled shapes). */
#include "fpa_kernel_common.h"

#define FPA_LAUNCH_CASE_sub(n, K_, L_) \
  if (K == K_ && L == L_) return fpa::launch_shape<K_, L_, 2>(*a, st);
FPA_DEFINE_LAUNCHER(sub, 2)
